"""Taichi-field-compatible views over device tensors.

The reference exposes every solver field as a Taichi field with
``to_numpy()``, ``from_numpy()``, ``fill()`` and ``x[None]`` for 0-d fields
(iceflow3d/simulator.py:78-141; callers and the checkpoint rely on exactly
these methods, checkpoint.py:49-56).  Here each field is a view over a
PyTorch-ROCm tensor whose physical layout is chosen for the GPU
(structure-of-arrays planes, int8 labels); the views convert to and from the
reference's logical layout (cell-major ``[nx, ny, nz, c]``, int32 labels).
"""

import numpy as np
import torch


class DeviceField:
    """A grid or per-body field.

    ``tensor`` holds the physical layout.  ``components`` > 0 means the
    physical tensor is ``[c, *shape]`` (SoA) and the logical array is
    ``[*shape, c]`` (``matrix`` reshapes the trailing axis to 3x3).
    """

    def __init__(self, tensor, *, components=0, logical_dtype=None, matrix=False, on_write=None):
        self.t = tensor
        self.components = components
        self.logical_dtype = logical_dtype or _numpy_dtype(tensor.dtype)
        self.matrix = matrix
        self.on_write = on_write

    @property
    def shape(self):
        base = tuple(self.t.shape[1:]) if self.components else tuple(self.t.shape)
        if self.components:
            base = base + ((3, 3) if self.matrix else (self.components,))
        return base

    @property
    def dtype(self):
        return self.logical_dtype

    def to_numpy(self):
        array = self.t.detach().cpu().numpy()
        if self.components:
            array = np.moveaxis(array, 0, -1)
            if self.matrix:
                array = array.reshape(*array.shape[:-1], 3, 3)
        return np.ascontiguousarray(array.astype(self.logical_dtype, copy=False))

    def _physical(self, value):
        array = np.asarray(value)
        if self.components:
            if self.matrix:
                array = array.reshape(*array.shape[:-2], 9)
            array = np.moveaxis(array, -1, 0)
        return torch.from_numpy(np.ascontiguousarray(array).astype(_numpy_dtype(self.t.dtype), copy=False))

    def from_numpy(self, value):
        value = np.asarray(value)
        if value.shape != self.shape:
            value = np.broadcast_to(value, self.shape)
        self.t.copy_(self._physical(value))
        if self.on_write is not None:
            self.on_write()

    def fill(self, value):
        value = np.asarray(value, dtype=float)
        if value.ndim == 0:
            self.t.fill_(float(value))
            if self.on_write is not None:
                self.on_write()
        else:
            self.from_numpy(np.broadcast_to(value, self.shape))

    def __getitem__(self, key):
        return self.to_numpy()[key]

    def __setitem__(self, key, value):
        array = self.to_numpy()
        array[key] = value
        self.from_numpy(array)


class DerivedField:
    """A read-only field computed on demand (writes are ignored because the
    value is a function of other state, e.g. capillary stress of the phase
    gradient)."""

    def __init__(self, compute, shape, dtype=np.float32):
        self._compute = compute
        self.shape = shape
        self.dtype = dtype

    def to_numpy(self):
        return self._compute()

    def from_numpy(self, value):
        pass

    def fill(self, value):
        pass

    def __getitem__(self, key):
        return self.to_numpy()[key]


class ControlScalar:
    """A 0-d field stored in the device control block."""

    def __init__(self, sim, name, dtype):
        self.sim, self.name, self.dtype = sim, name, dtype

    def __getitem__(self, key):
        value = getattr(self.sim._read_control(), self.name)
        if hasattr(value, "__len__"):
            return np.array(list(value), dtype=self.dtype)
        return self.dtype(value)

    def __setitem__(self, key, value):
        control = self.sim._read_control()
        current = getattr(control, self.name)
        if hasattr(current, "__len__"):
            for i, v in enumerate(np.asarray(value).ravel()):
                current[i] = v.item()
        else:
            setattr(control, self.name, np.asarray(value, dtype=self.dtype).item())
        self.sim._write_control(control)

    def to_numpy(self):
        return np.asarray(self[None], dtype=self.dtype)

    def from_numpy(self, value):
        self[None] = value

    def fill(self, value):
        self[None] = value


class HostScalar:
    """A 0-d field kept on the host (e.g. the AA parity bit)."""

    def __init__(self, obj, attr, dtype=np.int32):
        self.obj, self.attr, self.dtype = obj, attr, dtype

    def __getitem__(self, key):
        return self.dtype(getattr(self.obj, self.attr))

    def __setitem__(self, key, value):
        setattr(self.obj, self.attr, int(np.asarray(value).item()))

    def to_numpy(self):
        return np.asarray(self[None], dtype=self.dtype)

    def from_numpy(self, value):
        self[None] = value

    def fill(self, value):
        self[None] = value


def _numpy_dtype(torch_dtype):
    return {
        torch.float32: np.float32,
        torch.float64: np.float64,
        torch.int32: np.int32,
        torch.int8: np.int8,
        torch.uint8: np.uint8,
    }[torch_dtype]
