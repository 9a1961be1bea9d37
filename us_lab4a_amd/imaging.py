"""Host-side helpers of the STA path: grid points, apodization windows, array
geometry, and the acquisition interface.

Each follows the reference routine it names; they run on the host once per
geometry (not per frame) and carry no per-pixel work.
"""
import math

import numpy as np

# Matrix<XYZValueFactor<float>> element (XYZValueFactor.h:23-32): x, y, z,
# value, factor; factor defaults to 1.
XYZ_VALUE_FACTOR = np.dtype([("x", np.float32), ("y", np.float32), ("z", np.float32),
                             ("value", np.float32), ("factor", np.float32)])


def fill_sequence_start_end_size(start, end, size):
    """Util::fillSequenceFromStartToEndWithSize (Util.h:326-345): step in double."""
    if size < 1:
        raise ValueError("The size must be >= 1.")
    if size == 1:
        return np.array([start], dtype=np.float64)
    step = (float(end) - float(start)) / (size - 1)
    return float(start) + np.arange(size, dtype=np.float64) * step


def fill_sequence_start_end_max_step(start, end, step):
    """Util::fillSequenceFromStartToEndWithMaximumStep (Util.h:266-300)."""
    if step == 0:
        raise ValueError("The step must not be zero.")
    num_steps = math.ceil((float(end) - float(start)) / step)
    n = 1 + int(num_steps)
    return fill_sequence_start_end_size(start, end, n)


def rectangular_grid(min_x, max_x, n_x, min_z, max_z, n_z, origin_x=0.0, origin_z=0.0):
    """x-z grid as Matrix<XYZValueFactor<float>>[n1 = x columns][n2 = z rows]
    (ImageGrid.h:137-143 layout, z fastest) with explicit point counts."""
    xs = fill_sequence_start_end_size(min_x, max_x, n_x) + origin_x
    zs = fill_sequence_start_end_size(min_z, max_z, n_z) + origin_z
    g = np.zeros((n_x, n_z), dtype=XYZ_VALUE_FACTOR)
    g["x"] = xs[:, None].astype(np.float32)
    g["z"] = zs[None, :].astype(np.float32)
    g["factor"] = 1.0
    return g


def rectangular_grid_lambda(min_x, max_x, min_z, max_z, lambda_, step_div, origin_x=0.0, origin_z=0.0):
    """ImageGrid::getRectangularGrid (ImageGrid.h:120-143): step = lambda/step_div."""
    max_step = lambda_ / step_div
    xs = fill_sequence_start_end_max_step(min_x, max_x, max_step) + origin_x
    zs = fill_sequence_start_end_max_step(min_z, max_z, max_step) + origin_z
    g = np.zeros((xs.size, zs.size), dtype=XYZ_VALUE_FACTOR)
    g["x"] = xs[:, None].astype(np.float32)
    g["z"] = zs[None, :].astype(np.float32)
    g["factor"] = 1.0
    return g


def grid_xz(grid):
    """[n1][n2][2] float32 (x, z) view used by the C ABI."""
    return np.ascontiguousarray(np.stack([grid["x"], grid["z"]], axis=-1), dtype=np.float32)


def element_x_centered_2d(num_elements, pitch):
    """ArrayGeometry::getElementXCentered2D (ArrayGeometry.h:37-49), float."""
    if num_elements == 0:
        raise ValueError("The number of elements is zero.")
    pitch = np.float32(pitch)
    xc = np.float32(np.float32((num_elements - 1) * pitch) / np.float32(2))
    return (np.arange(num_elements, dtype=np.float32) * pitch - xc).astype(np.float32)


def element_positions_2d_array(pitch_x, n_x, offset_x, pitch_y, n_y, offset_y):
    """ArrayUtil::calculateElementPositions (ArrayUtil.h:51-69): [n_y*n_x][2] (x, y)."""
    half_w = (n_x - 1) * 0.5 * pitch_x
    half_h = (n_y - 1) * 0.5 * pitch_y
    pos = np.empty((n_y * n_x, 2), dtype=np.float32)
    for iy in range(n_y):
        for ix in range(n_x):
            pos[iy * n_x + ix, 0] = ix * pitch_x - half_w + offset_x
            pos[iy * n_x + ix, 1] = iy * pitch_y - half_h + offset_y
    return pos


def window(name, n):
    """WindowFunction::get (WindowFunction.h:237-289) for the fixed-shape windows."""
    if n < 2:
        raise ValueError("The window function size must be >= 2.")
    i = np.arange(n, dtype=np.float64)
    c1 = 1.0 / (n - 1.0)
    c3 = 1.0 / (n + 1.0)
    if name == "rectangular":
        w = np.ones(n)
    elif name == "hanning":
        w = 0.5 - 0.5 * np.cos(2.0 * math.pi * (i * c1))
    elif name == "hanning2":
        w = 0.5 - 0.5 * np.cos(2.0 * math.pi * ((i + 1.0) * c3))
    elif name == "hamming":
        w = 0.54 - 0.46 * np.cos(2.0 * math.pi * (i * c1))
    elif name == "blackman":
        c2 = i * c1
        w = 0.42 - 0.5 * np.cos(2.0 * math.pi * c2) + 0.08 * np.cos(4.0 * math.pi * c2)
    elif name == "blackman2":
        c2 = (i + 1.0) * c3
        w = 0.42 - 0.5 * np.cos(2.0 * math.pi * c2) + 0.08 * np.cos(4.0 * math.pi * c2)
    else:
        raise ValueError(f"Invalid window function name: {name}.")
    return w.astype(np.float32)


def reset_value_factor(grid):
    """Util::resetValueFactor (Util.h:990-999)."""
    grid["value"] = 0.0
    grid["factor"] = 1.0


def apply_factor_to_value(grid):
    """Util::applyFactorToValue (Util.h:1052-1060): the coherence-factor image."""
    grid["value"] *= grid["factor"]
    grid["factor"] = 1.0


class STAAcquisition:
    """Echo source interface (STAAcquisition.h:71-86): execute(tx) returns the
    [rx][samples] float32 matrix of one transmit."""

    def prepare(self, base_element):
        pass

    def execute(self, tx_elem):
        raise NotImplementedError


class ArrayAcquisition(STAAcquisition):
    """In-memory acquisition over a [ntx][nrx][ns] array (tx indexed from first_tx)."""

    def __init__(self, echoes, first_tx=0):
        self.echoes = np.ascontiguousarray(echoes, dtype=np.float32)
        self.first_tx = first_tx

    def execute(self, tx_elem):
        return self.echoes[tx_elem - self.first_tx]
