"""Raw binary fields with XDMF metadata, the file pair the reference writes
for visualisation (`util_io.py:8-144`): `<name>.raw` holds the C-order
values (axes (z, y, x), little-endian FP64 or FP32) and `<name>.xmf` an XDMF
2.0 description of a uniform 3D co-rectilinear grid whose one scalar
attribute is a hyperslab of that file. 2D fields are written as one z
layer. Host-side I/O (device tensors are copied to the host first).

`read_raw` takes the value type from the XMF `Precision` attribute (the
reference always reads FP64, which misreads its own FP32 files).
"""

import os
import re

import numpy as np

__all__ = ['parse_raw_xmf', 'read_raw', 'write_raw_xmf', 'write_raw_with_xmf']


def _binary_item(xmfpath):
    """(dimensions, precision, file name) of the binary DataItem."""
    with open(xmfpath) as f:
        text = ' '.join(f.read().split())
    items = re.findall(r'<DataItem([^>]*)>\s*([^<\s]*)\s*</DataItem>', text)
    for attrs, body in items:
        if 'Format="Binary"' in attrs:
            dims = re.search(r'Dimensions="([0-9 ]+)"', attrs).group(1)
            prec = re.search(r'Precision="([0-9]+)"', attrs)
            return (tuple(int(x) for x in dims.split()),
                    int(prec.group(1)) if prec else 8, body)
    raise ValueError('{}: no binary DataItem'.format(xmfpath))


def parse_raw_xmf(xmfpath):
    """(shape, raw path) of the field described by `xmfpath`; the raw path
    is relative to the XMF file's folder."""
    shape, _, raw = _binary_item(xmfpath)
    return shape, os.path.join(os.path.dirname(xmfpath), raw)


def read_raw(xmfpath):
    """The array of a raw file described by `xmfpath`."""
    shape, prec, raw = _binary_item(xmfpath)
    dtype = np.float32 if prec == 4 else np.float64
    path = os.path.join(os.path.dirname(xmfpath), raw)
    return np.fromfile(path, dtype=dtype).reshape(shape)


def write_raw_xmf(xmfpath, rawpath, count, spacing=(1, 1, 1), name='data',
                  precision=8, cell=True):
    """XDMF metadata for a raw file of shape `count` = (nz, ny, nx):
    origin 0, the given cell `spacing` (x, y, z order reversed in the file
    like the axes), cell- or node-centred scalar `name`."""
    dim = 3
    count = [int(c) for c in count]

    def rev(v):
        return ' '.join(str(x) for x in reversed(list(v)))

    def fwd(v):
        return ' '.join(str(x) for x in v)

    nodes = [c + 1 for c in count] if cell else count
    kind = 'Double' if precision == 8 else 'Float'
    lines = [
        '<?xml version="1.0" ?>',
        '<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd" []>',
        '<Xdmf Version="2.0">',
        ' <Domain>',
        '   <Grid Name="mesh" GridType="Uniform">',
        '     <Topology TopologyType="{}DCORECTMesh" Dimensions="{}"/>'.format(
            dim, fwd(nodes)),
        '     <Geometry GeometryType="ORIGIN_DXDYDZ">',
        '       <DataItem Name="Origin" Dimensions="{}" NumberType="Float" '
        'Precision="8" Format="XML">'.format(dim),
        '         ' + rev([0] * dim),
        '       </DataItem>',
        '       <DataItem Name="Spacing" Dimensions="{}" NumberType="Float" '
        'Precision="8" Format="XML">'.format(dim),
        '         ' + rev(spacing),
        '       </DataItem>',
        '     </Geometry>',
        '     <Attribute Name="{}" AttributeType="Scalar" Center="{}">'.format(
            name, 'Cell' if cell else 'Node'),
        '       <DataItem ItemType="HyperSlab" Dimensions="{}" '
        'Type="HyperSlab">'.format(fwd(count)),
        '           <DataItem Dimensions="3 {}" Format="XML">'.format(dim),
        '             ' + rev([0] * dim),
        '             ' + rev([1] * dim),
        '             ' + fwd(count),
        '           </DataItem>',
        '           <DataItem Dimensions="{}" Seek="0" Precision="{}" '
        'NumberType="{}" Format="Binary">'.format(fwd(count), precision,
                                                   kind),
        '             ' + str(rawpath),
        '           </DataItem>',
        '       </DataItem>',
        '     </Attribute>',
        '   </Grid>',
        ' </Domain>',
        '</Xdmf>',
    ]
    with open(xmfpath, 'w') as f:
        f.write('\n'.join(lines) + '\n')


def write_raw_with_xmf(u, xmfpath, rawpath=None, spacing=(1, 1, 1),
                       cell=True, name='data'):
    """Writes `u` (host array or device tensor, 2D or 3D) as raw data next
    to its XMF description; returns xmfpath."""
    if hasattr(u, 'detach'):
        u = u.detach().cpu().numpy()
    u = np.asarray(u)
    if u.dtype not in (np.float32, np.float64):
        u = u.astype(np.float64)
    if u.ndim != 3:
        u = u.reshape((1, ) + u.shape)
    spacing = list(spacing)
    if len(spacing) != 3:
        spacing = spacing + [min(spacing)]
    precision = 4 if u.dtype == np.float32 else 8
    if rawpath is None:
        rawpath = os.path.splitext(xmfpath)[0] + '.raw'
    write_raw_xmf(xmfpath, os.path.basename(rawpath)
                  if os.path.dirname(rawpath) == os.path.dirname(xmfpath)
                  else rawpath, u.shape, spacing, name, precision, cell)
    np.ascontiguousarray(u).tofile(rawpath)
    return xmfpath
