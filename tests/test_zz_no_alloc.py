"""opir_ext_prod_mux_device_noalloc: the allocation-free, device-pointer form
of opir_ext_prod_mux (PirServer::ext_prod_mux, src/server.cpp:652-722),
checked bit-exact against the existing entry point on two calls in a row
(the second with a smaller selector gadget), for both schemes. Device
buffers are allocated by the test through the HIP runtime the library links;
the library itself allocates nothing.
"""
import ctypes
import sys

import numpy as np
import pytest

from conftest import scheme_from_oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


_HIP = None


def _hip():
    """The HIP runtime the library itself links (so both share one runtime
    and one device context); used here only to hold the test's operands."""
    global _HIP
    if _HIP is None:
        import os
        from onionpirv2_amd import gpu
        gpu.load_library()
        path = "/opt/rocm/lib/libamdhip64.so.7"
        _HIP = ctypes.CDLL(path if os.path.exists(path) else "libamdhip64.so")
        _HIP.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _HIP.hipFree.argtypes = [ctypes.c_void_p]
        _HIP.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _HIP.hipDeviceSynchronize.argtypes = []
    return _HIP


class _Dev:
    """A device copy of a uint64 array (hipMalloc'd by the test, not the library)."""

    def __init__(self, a):
        self.a = np.ascontiguousarray(a, dtype=np.uint64)
        self.nbytes = self.a.nbytes
        self.p = ctypes.c_void_p()
        assert _hip().hipMalloc(ctypes.byref(self.p), self.nbytes) == 0
        assert _hip().hipMemcpy(self.p, self.a.ctypes.data, self.nbytes, 1) == 0  # host to device

    def get(self):
        out = np.empty_like(self.a)
        assert _hip().hipMemcpy(out.ctypes.data, self.p, self.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        if self.p:
            _hip().hipFree(self.p)
            self.p = ctypes.c_void_p()


def _noalloc(ctx, gsw, l, b, x, y, alias_x=False):
    dg, dx, dy = _Dev(gsw), _Dev(x), _Dev(y)
    dout = dx if alias_x else _Dev(np.zeros_like(x))
    try:
        assert _hip().hipDeviceSynchronize() == 0
        rc = ctx.lib.opir_ext_prod_mux_device_noalloc(ctx.h, dg.p, l, b, dx.p, dy.p, dout.p)
        ctx.check(rc)
        ctx.synchronize()
        out = dout.get()
        if not alias_x:
            assert np.array_equal(dx.get(), x) and np.array_equal(dy.get(), y), "inputs must stay unchanged"
        return out
    finally:
        for d in {id(v): v for v in (dg, dx, dy, dout)}.values():
            d.free()


def _random_ct(rng, mods, N):
    return np.concatenate([rng.integers(0, q, size=(2, N), dtype=np.uint64) for q in mods], axis=1)


def test_noalloc_mux_matches_ext_prod_mux_composite(O, comp_params):
    from onionpirv2_amd.gpu import GpuContext, InvalidArgument
    say("[gpu] no-alloc ext_prod_mux, composite scheme")
    p = comp_params
    Q = p.rns_mods[0]
    ctx = GpuContext(scheme_from_oracle(p))
    try:
        rng = np.random.default_rng(71)
        for l, b, alias in [(p.l_key, p.base_log2_key, False), (p.l_ep, p.base_log2, True)]:
            gsw = rng.integers(0, Q, size=(2 * l, 2 * p.N), dtype=np.uint64)
            x, y = _random_ct(rng, [Q], p.N), _random_ct(rng, [Q], p.N)
            x[0, :3] = [0, Q - 1, 7]
            y[0, :3] = [Q - 1, 0, 7]
            want = ctx.ext_prod_mux(gsw, l, b, x, y)
            assert np.array_equal(_noalloc(ctx, gsw, l, b, x, y, alias), want), (l, b)
        with pytest.raises(InvalidArgument):
            _noalloc(ctx, np.zeros((34, 2 * p.N), np.uint64), 17, 4, x, y)
    finally:
        ctx.close()


def test_noalloc_mux_matches_ext_prod_mux_two_primes(O):
    from onionpirv2_amd.gpu import GpuContext
    say("[gpu] no-alloc ext_prod_mux, two-prime scheme")
    p = O.PirParams("CONFIG_N4096_K2_MP")
    q0, q1 = p.rns_mods
    sch = dict(poly_degree=p.N, composite=0, q1=q0, q2=q1, ntt_root=O.ntt_root(p.N, q0),
               ntt_root2=O.ntt_root(p.N, q1), plain_mod=p.plain_mod, small_q=p.small_q, l_ep=p.l_ep,
               base_log2=p.base_log2, l_key=p.l_key, base_log2_key=p.base_log2_key, l_ks=p.l_ks,
               bv_base_log2=p.bv_base_log2)
    ctx = GpuContext(sch)
    try:
        rng = np.random.default_rng(72)
        for l, b in [(p.l_key, p.base_log2_key), (p.l_ep, p.base_log2)]:
            gsw = np.concatenate([rng.integers(0, q, size=(2 * l, 2, p.N), dtype=np.uint64) for q in (q0, q1)],
                                 axis=2).reshape(2 * l, 4 * p.N)
            x, y = _random_ct(rng, [q0, q1], p.N), _random_ct(rng, [q0, q1], p.N)
            want = ctx.ext_prod_mux(gsw, l, b, x, y)
            assert np.array_equal(_noalloc(ctx, gsw, l, b, x, y), want), (l, b)
    finally:
        ctx.close()


def test_noalloc_mux_matches_ext_prod_mux_single_prime(O):
    from onionpirv2_amd.gpu import GpuContext
    say("[gpu] no-alloc ext_prod_mux, single-prime scheme")
    p = O.PirParams("CONFIG_N2048_K1")
    (q,) = p.rns_mods
    sch = dict(poly_degree=p.N, composite=0, q1=q, q2=0, ntt_root=O.ntt_root(p.N, q), plain_mod=p.plain_mod,
               small_q=p.small_q, l_ep=p.l_ep, base_log2=p.base_log2, l_key=p.l_key, base_log2_key=p.base_log2_key,
               l_ks=p.l_ks, bv_base_log2=p.bv_base_log2)
    ctx = GpuContext(sch)
    try:
        rng = np.random.default_rng(73)
        for l, b in [(p.l_key, p.base_log2_key), (p.l_ep, p.base_log2)]:
            gsw = rng.integers(0, q, size=(2 * l, 2 * p.N), dtype=np.uint64)
            x, y = _random_ct(rng, [q], p.N), _random_ct(rng, [q], p.N)
            x[0, :3] = [0, q - 1, 7]
            y[0, :3] = [q - 1, 0, 7]
            want = ctx.ext_prod_mux(gsw, l, b, x, y)
            assert np.array_equal(_noalloc(ctx, gsw, l, b, x, y), want), (l, b)
    finally:
        ctx.close()
