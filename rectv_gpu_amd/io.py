"""Output of reconstructions as TIFF stacks, the way the reference's examples save
them (examples/foam/foam_example.py:103-104, examples/tao/tao.py:177-178:
dxchange.write_tiff_stack(rtv[k], 'rec_tv/rec_' + str(k)) for every slice k).

write_tiff_stack follows dxchange's naming: one 32-bit float TIFF per index along
`axis`, named <fname>_<index, zero-padded to `digit` places>.tiff.  Written with
Pillow (dxchange and tifffile are not installed in this image).  HDF5 input
(/exchange/data, /exchange/theta, examples/tao/tao.py:46-52) needs h5py, which is
not installed either, so it is not provided.
"""
import os

import numpy as np


def _name(fname, index, digit):
    body = fname[:-5] if fname.endswith(".tiff") else (fname[:-4] if fname.endswith(".tif") else fname)
    return "%s_%0*d.tiff" % (body, digit, index)


def write_tiff_stack(data, fname="tmp/data", axis=0, digit=5, start=0, overwrite=False):
    """Write data as a stack of float32 TIFF images, one per index along `axis`
    (files <fname>_<start+i>.tiff).  Existing files raise FileExistsError unless
    overwrite is set.  Returns the list of file names."""
    from PIL import Image
    a = np.moveaxis(np.asarray(data, dtype=np.float32), axis, 0)
    if a.ndim != 3:
        raise ValueError("data must be 3-D")
    d = os.path.dirname(fname)
    if d:
        os.makedirs(d, exist_ok=True)
    names = []
    for i in range(a.shape[0]):
        nm = _name(fname, start + i, digit)
        if os.path.exists(nm) and not overwrite:
            raise FileExistsError(nm)
        Image.fromarray(np.ascontiguousarray(a[i]), mode="F").save(nm, format="TIFF")
        names.append(nm)
    return names


def read_tiff_stack(fname, count, digit=5, start=0):
    """Read count images written by write_tiff_stack back as [count, rows, cols] float32."""
    from PIL import Image
    return np.stack([np.asarray(Image.open(_name(fname, start + i, digit)), dtype=np.float32) for i in range(count)])


def save_reconstruction(rtv, prefix="rec_tv/rec", overwrite=True):
    """The examples' loop: for each slice k of rtv [nz, m, n, n], the m time frames as
    the stack <prefix>_<k>_<frame>.tiff."""
    names = []
    for k in range(rtv.shape[0]):
        names += write_tiff_stack(rtv[k], "%s_%d" % (prefix, k), overwrite=overwrite)
    return names
