"""Running median of a video stream on the GPU, mirroring
utilities/vmedian.py:6-121.

A tree of 3-frame median buffers of depth ``order``: each level's median
(lmz_median3_u8, pixel by pixel) updates once per three updates of the
level below, so a tree of depth ``order`` has seen 3**(order + 1) frames at
its first output.  The buffers are device memory; frames may be numpy
arrays or device tensors, and ``get()`` answers in the kind last given.
"""
from __future__ import annotations

import numpy as np

from .. import _lmz
from .normalizer import is_device_tensor


class VMedian:
    """Running median of a video stream (utilities/vmedian.py:6-121)."""

    def __init__(self, order: int = 0, shape: tuple | None = None) -> None:
        self.child = None
        self._shape = None
        self.index = 0
        self._initialized = False
        self._cycled = False
        self._host = True
        self._data = None
        self.order = order
        self.shape = shape

    def filter(self, data):
        """Add ``data`` and return the current median estimate."""
        self.add(data)
        return self.get()

    def get(self, reshape: bool = True):
        """The current median image (vmedian.py:35-48): a new array or
        tensor each call, which later frames do not change (the running
        median is updated in place)."""
        d = self._data
        if reshape:
            d = d.reshape(self.shape)
        return d.cpu().numpy() if self._host else d.clone()

    def add(self, data) -> None:
        """Include a new frame (vmedian.py:50-77)."""
        t = _lmz.torch()
        self._host = not is_device_tensor(data)
        frame = self._as_u8(data)
        if tuple(frame.shape) != self.shape:
            self._data = frame.reshape(-1).clone()
            self.shape = tuple(frame.shape)
        if self.order == 0:
            self.buffer[self.index].copy_(frame.reshape(-1))
            self.index += 1
        else:
            child = self.child
            child.add(data)
            if child.initialized:
                self.buffer[self.index].copy_(child._data)
            if child.cycled:
                self.index += 1
        if self.index == 3:
            self.index = 0
            if self._data is None or self._data.numel() != self.buffer.shape[1]:
                self._data = t.empty(self.buffer.shape[1], dtype=t.uint8,
                                     device=self.buffer.device)
            _lmz.check(_lmz.load_library().lmz_median3_u8(
                _lmz.ptr(self.buffer), self.buffer.shape[1],
                _lmz.ptr(self._data), _lmz.stream_handle()))
            self._initialized = True
            self._cycled = True
        else:
            self._cycled = False

    def reset(self) -> None:
        """Reset the filter state (vmedian.py:79-84)."""
        self._initialized = False
        self._cycled = False
        if self.order > 0:
            self.child.reset()

    @property
    def initialized(self) -> bool:
        return self._initialized

    @property
    def cycled(self) -> bool:
        return self._cycled

    @property
    def shape(self):
        return self._shape

    @shape.setter
    def shape(self, shape) -> None:
        """Reallocate the buffers and restart (vmedian.py:100-110)."""
        self._shape = None if shape is None else tuple(int(v) for v in shape)
        if shape is None:
            return
        if self.child is not None:
            self.child.shape = shape
        t = _lmz.torch()
        npts = int(np.prod(shape))
        self.buffer = t.zeros((3, npts), dtype=t.uint8, device=_lmz.device())
        self.index = 0
        self._initialized = False

    @property
    def order(self) -> int:
        return self._order

    @order.setter
    def order(self, order: int) -> None:
        self._order = int(np.clip(order, 0, 10))
        if self._order > 0:
            self.child = VMedian(order=self._order - 1, shape=self.shape)
        self._initialized = False

    @staticmethod
    def _as_u8(data):
        """data.astype(np.uint8) on the device."""
        t = _lmz.torch()
        if is_device_tensor(data):
            return data.to(t.uint8).contiguous()
        arr = np.asarray(data).astype(np.uint8)
        return t.from_numpy(np.ascontiguousarray(arr)).to(_lmz.device())


vmedian = VMedian  # backward-compatible alias (vmedian.py:124)
