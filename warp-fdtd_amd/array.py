"""Device arrays with the small part of the wp.array surface that callers of
the reference solver use (``numpy()``, ``zero_()``, ``assign()``, ``capacity``,
``ptr``), backed by a contiguous torch tensor on the HIP device."""

import numpy as np
import torch

_TORCH = {
    np.dtype(np.float32): torch.float32,
    np.dtype(np.float64): torch.float64,
    np.dtype(np.int32): torch.int32,
    np.dtype(np.int64): torch.int64,
}


def torch_dtype(dtype):
    return _TORCH[np.dtype(dtype)]


class Array:
    """A device array; ``tensor`` is the torch view the kernels write into."""

    __slots__ = ("tensor",)

    def __init__(self, tensor):
        if not tensor.is_contiguous():
            raise ValueError("Array needs a contiguous tensor")
        self.tensor = tensor

    @classmethod
    def zeros(cls, shape, dtype=np.float32, device="cuda:0"):
        return cls(torch.zeros(shape, dtype=torch_dtype(dtype), device=device))

    @classmethod
    def empty(cls, shape, dtype=np.float32, device="cuda:0"):
        return cls(torch.empty(shape, dtype=torch_dtype(dtype), device=device))

    @classmethod
    def from_numpy(cls, a, dtype=None, device="cuda:0"):
        a = np.ascontiguousarray(a, dtype=dtype)
        return cls(torch.from_numpy(a.copy()).to(device))

    @property
    def shape(self):
        return tuple(self.tensor.shape)

    @property
    def ptr(self):
        return self.tensor.data_ptr()

    @property
    def capacity(self):
        return self.tensor.numel() * self.tensor.element_size()

    @property
    def size(self):
        return self.tensor.numel()

    def numpy(self):
        return self.tensor.detach().cpu().numpy()

    def zero_(self):
        self.tensor.zero_()
        return self

    def assign(self, values):
        """Overwrite in place (pointers captured by graphs stay valid)."""
        src = torch.as_tensor(np.asarray(values)).to(self.tensor.dtype)
        self.tensor.copy_(src.reshape(self.tensor.shape))
        return self

    def __len__(self):
        return self.tensor.shape[0]

    def __getitem__(self, index):
        return Array(self.tensor[index])

    def __repr__(self):
        return f"Array(shape={self.shape}, dtype={self.tensor.dtype}, device={self.tensor.device})"
