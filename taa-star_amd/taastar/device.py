"""Device-side helpers: frame buffers owned by torch tensors, calls through the C ABI."""
import ctypes as C

import numpy as np
import torch

from . import (FMT_RGBA16F, INSPECT_WORDS, VEL_RG16F, VEL_RG32F, VEL_RGBA32F, FrameImages, Reproject, check, lib)


def _stream_ptr(stream):
    if stream is None:
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


def to_dev(a, device="cuda"):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def color_dtype(fmt):
    return torch.float16 if fmt == FMT_RGBA16F else torch.float32


def vel_shape_dtype(vfmt):
    ch = 2 if vfmt in (VEL_RG32F, VEL_RG16F) else 4
    dt = torch.float32 if vfmt in (VEL_RG32F, VEL_RGBA32F) else torch.float16
    return ch, dt


class FrameBuffers:
    """Every image one resolve touches, as device tensors [H, W, C]. input_size (Wl, Hl): the
    render size of colour, depth, velocity and previous depth when it differs from the output."""

    def __init__(self, width, height, fmt, vfmt, debug=False, inspect=False, device="cuda", input_size=None):
        self.W, self.H, self.fmt, self.vfmt = width, height, fmt, vfmt
        self.Wl, self.Hl = input_size if input_size is not None else (width, height)
        cd = color_dtype(fmt)
        vc, vd = vel_shape_dtype(vfmt)
        z = lambda *s, dtype: torch.zeros(*s, dtype=dtype, device=device)
        self.color = z(self.Hl, self.Wl, 4, dtype=cd)
        self.depth = z(self.Hl, self.Wl, dtype=torch.float32)
        self.velocity = z(self.Hl, self.Wl, vc, dtype=vd)
        self.history = [z(height, width, 4, dtype=cd), z(height, width, 4, dtype=cd)]
        self.prev_depth = z(self.Hl, self.Wl, dtype=torch.float32)
        self.screen = z(height, width, 4, dtype=cd)
        self.debug = z(height, width, 4, dtype=torch.float16) if debug else None
        self.inspect = z(height, width, INSPECT_WORDS, dtype=torch.int32) if inspect else None
        self.k = 0  # frames resolved

    def upload(self, color, depth, velocity):
        """Store a frame's inputs (numpy arrays already in the storage dtype)."""
        self.prev_depth.copy_(self.depth)
        self.color.copy_(torch.from_numpy(np.ascontiguousarray(color)))
        self.depth.copy_(torch.from_numpy(np.ascontiguousarray(depth)))
        self.velocity.copy_(torch.from_numpy(np.ascontiguousarray(velocity)))

    @property
    def hist_in(self):
        return self.history[(self.k + 1) & 1]

    @property
    def hist_out(self):
        return self.history[self.k & 1]

    def images(self, screen=None):
        img = FrameImages()
        img.color = self.color.data_ptr()
        img.depth = self.depth.data_ptr()
        img.history = self.hist_in.data_ptr()
        img.history_depth = self.prev_depth.data_ptr()
        img.result_screen = (self.screen if screen is None else screen).data_ptr()
        img.debug = self.debug.data_ptr() if self.debug is not None else None
        img.velocity = self.velocity.data_ptr()
        img.result_history = self.hist_out.data_ptr()
        img.inspect = self.inspect.data_ptr() if self.inspect is not None else None
        img.width, img.height, img.pitch = self.W, self.H, self.W
        if (self.Wl, self.Hl) != (self.W, self.H):
            img.input_width, img.input_height, img.input_pitch = self.Wl, self.Hl, self.Wl
        img.color_format, img.velocity_format = self.fmt, self.vfmt
        return img

    def resolve(self, uniforms, cas=None, arith=0, stream=None, screen=None):
        img = self.images(screen)
        check(lib().taa_resolve_frame(C.byref(uniforms), C.byref(img), C.byref(cas) if cas is not None else None,
                                      arith, _stream_ptr(stream)))
        self.k += 1


def cas_apply(src, dst, fmt, cas, stream=None):
    H, W = src.shape[0], src.shape[1]
    check(lib().taa_cas_apply(C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()), W, H, W, fmt, C.byref(cas),
                              _stream_ptr(stream)))


def velocity_from_depth(reproject, depth, velocity, vfmt, ids=None, objects=None, width=None, stream=None):
    """taa_velocity_from_depth on torch tensors. depth [H, P] float32, velocity [H, P, ch] in vfmt,
    ids [H, P] int32/uint32 or None, objects a device float32 tensor [n, 16] or None; width <= P
    (default P) is the image width, P the pitch."""
    H, P = depth.shape[0], depth.shape[1]
    r = Reproject.from_buffer_copy(reproject)
    r.object_prev_from_current = None if objects is None else objects.data_ptr()
    check(lib().taa_velocity_from_depth(C.byref(r), C.c_void_p(depth.data_ptr()),
                                        None if ids is None else C.c_void_p(ids.data_ptr()),
                                        C.c_void_p(velocity.data_ptr()), vfmt, width or P, H, P, _stream_ptr(stream)))
