"""Oracle restatement of the volume creator's shape painters (TEST
INFRASTRUCTURE ONLY): src/kernels/DrawGGEMSBox.cl:82-115,
DrawGGEMSTube.cl:77-108, DrawGGEMSSphere.cl:78-102, in numpy FP32 with
the kernels' order of operations and their voxel-index arithmetic."""

import numpy as np

F = np.float32
MET = {"MET_CHAR": np.int8, "MET_UCHAR": np.uint8, "MET_SHORT": np.int16, "MET_USHORT": np.uint16,
       "MET_INT": np.int32, "MET_UINT": np.uint32, "MET_FLOAT": np.float32}


def voxel_indices(dims):
    """The kernels' decomposition of the voxel id; y takes the id modulo
    dims.x * dims.x (right for square x/y only)."""
    dx, dy, dz = (int(d) for d in dims)
    gid = np.arange(dx * dy * dz, dtype=np.int64)
    iy = (gid % (dx * dx)) // dx
    ix = (gid % (dx * dy)) - iy * dx
    iz = gid // (dx * dy)
    return ix, iy, iz


def _centre(es, n, i, pos):
    h = F(es) / F(2.0)
    k = (F(1.0) - F(n)) + F(2.0) * i.astype(F)
    return (h * k).astype(F) - F(pos)


def draw(volume, dims, element_sizes, shape, position, label, size):
    """volume: flat array (x fastest) of the element type; shape 'box'
    (size = height, width, depth), 'tube' (height, radius_x, radius_y) or
    'sphere' (radius,)."""
    ix, iy, iz = voxel_indices(dims)
    x = _centre(element_sizes[0], dims[0], ix, position[0])
    y = _centre(element_sizes[1], dims[1], iy, position[1])
    z = _centre(element_sizes[2], dims[2], iz, position[2])
    if shape == "box":
        hh, hw, hd = F(size[0]) / F(2.0), F(size[1]) / F(2.0), F(size[2]) / F(2.0)
        m = (z <= hd) & (z >= -hd) & (y <= hh) & (y >= -hh) & (x <= hw) & (x >= -hw)
    elif shape == "tube":
        h2 = F(size[0]) / F(2.0)
        rx2, ry2 = F(size[1]) * F(size[1]), F(size[2]) * F(size[2])
        m = (z <= h2) & (z >= -h2) & (((x * x) / rx2 + (y * y) / ry2) <= F(1.0))
    else:
        m = ((x * x + y * y) + z * z) <= F(size[0]) * F(size[0])
    volume[m] = np.array(F(label)).astype(volume.dtype)
    return volume
