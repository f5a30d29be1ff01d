"""Result reuse on device-resident arrays.

The timestep advects (u,v,w) and then (uo,vo,wo), where uo..wo are cpy_
copies of the previous u..w (srcs/main.f90:532-533, 686-688), and asks for
the unit normals of the same gradients at four staggerings in a row
(srcs/main.f90:549-558).  The library recognises through content versions
that a request repeats an earlier computation (or one computed alongside
it) and swaps in the parked result.  These tests drive those sequences
through the C ABI on registered host arrays and check every result bit for
bit against the oracle, including the cases where an input changed and the
result must be recomputed.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import felharness as H
from felharness import D, I, Grid

pytestmark = pytest.mark.gpu

NID = np.full(6, -1, dtype=np.int32)


def same_bits(a, b):
    """Bitwise equality (tells -0.0 from +0.0)."""
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


class Registered:
    """Registers numpy arrays with the library for the duration of a test."""

    def __init__(self):
        self.arrays = []

    def add(self, a):
        lib = H.gpu()
        n = ctypes.c_int64(a.nbytes)
        up = ctypes.c_int(1)
        lib.felgpu_register_(ctypes.c_void_p(a.ctypes.data), ctypes.byref(n), ctypes.byref(up))
        self.arrays.append(a)
        return a

    def download(self, a):
        H.gpu().felgpu_download_(ctypes.c_void_p(a.ctypes.data))
        return a

    def close(self):
        lib = H.gpu()
        for a in self.arrays:
            lib.felgpu_unregister_(ctypes.c_void_p(a.ctypes.data))
        self.arrays = []


def weno_oracle(g, u, v, w):
    out = [g.zeros() for _ in range(3)]
    H.call_oracle("cal_advu_weno5", I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy),
                  D(1 / g.dz), u.copy(), v.copy(), w.copy(), *out)
    return out


def test_advection_of_copied_inputs_is_reused_exactly():
    g = Grid(32, 16, 16, 0.1, 0.05, 0.1)
    r = np.random.default_rng(40)
    reg = Registered()
    try:
        u, v, w = (reg.add(r.standard_normal(g.shape)) for _ in range(3))
        uo, vo, wo = (reg.add(g.zeros()) for _ in range(3))
        x = reg.add(r.standard_normal(g.shape))
        adv = [reg.add(g.zeros()) for _ in range(3)]
        advo = [reg.add(g.zeros()) for _ in range(3)]
        dims = [I(g.ni), I(g.nj), I(g.nk)]
        inv = [D(1 / g.dx), D(1 / g.dy), D(1 / g.dz)]
        want0 = weno_oracle(g, u, v, w)
        u0 = u.copy()
        # step n: advect (u,v,w), then rotate: uo <- u, u <- x
        H.call_gpu("cal_advu_weno5", *dims, *inv, u, v, w, *adv)
        for a, b in ((u, uo), (v, vo), (w, wo)):
            H.call_gpu("cpy", *dims, a, b)
        H.call_gpu("cpy", *dims, x, u)
        # step n+1: advect the new (u,v,w), then (uo,vo,wo) -> parked result
        H.call_gpu("cal_advu_weno5", *dims, *inv, u, v, w, *adv)
        H.call_gpu("cal_advu_weno5", *dims, *inv, uo, vo, wo, *advo)
        for a in adv + advo:
            reg.download(a)
        want1 = weno_oracle(g, x, v, w)
        for got, want in zip(advo, want0):
            assert np.array_equal(got, want)
        for got, want in zip(adv, want1):
            assert np.array_equal(got, want)
        # an input written after the copy: recomputed, not reused
        H.call_gpu("cpy", *dims, x, vo)
        H.call_gpu("cal_advu_weno5", *dims, *inv, uo, vo, wo, *advo)
        for a in advo:
            reg.download(a)
        for got, want in zip(advo, weno_oracle(g, u0, x, w)):
            assert np.array_equal(got, want)
        # the same request twice in a row: the second is recomputed or reused,
        # either way identical
        H.call_gpu("cal_advu_weno5", *dims, *inv, uo, vo, wo, *advo)
        H.call_gpu("cal_advu_weno5", *dims, *inv, uo, vo, wo, *adv)
        for a in adv + advo:
            reg.download(a)
        for got, want in zip(adv, weno_oracle(g, u0, x, w)):
            assert np.array_equal(got, want)
    finally:
        reg.close()


def test_host_upload_invalidates_parked_result():
    g = Grid(16, 8, 16, 0.1, 0.1, 0.1)
    r = np.random.default_rng(41)
    reg = Registered()
    try:
        u, v, w = (reg.add(r.standard_normal(g.shape)) for _ in range(3))
        uo, vo, wo = (reg.add(g.zeros()) for _ in range(3))
        adv = [reg.add(g.zeros()) for _ in range(3)]
        advo = [reg.add(g.zeros()) for _ in range(3)]
        dims = [I(g.ni), I(g.nj), I(g.nk)]
        inv = [D(1 / g.dx), D(1 / g.dy), D(1 / g.dz)]
        H.call_gpu("cal_advu_weno5", *dims, *inv, u, v, w, *adv)
        for a, b in ((u, uo), (v, vo), (w, wo)):
            H.call_gpu("cpy", *dims, a, b)
        H.call_gpu("cal_advu_weno5", *dims, *inv, w, v, u, *adv)  # parks the first result
        # the host rewrites uo and uploads it: the parked result no longer applies
        reg.download(uo)
        uo[...] = r.standard_normal(g.shape)
        H.gpu().felgpu_upload_(ctypes.c_void_p(uo.ctypes.data))
        H.call_gpu("cal_advu_weno5", *dims, *inv, uo, vo, wo, *advo)
        for a in advo:
            reg.download(a)
        reg.download(vo)
        reg.download(wo)
        for got, want in zip(advo, weno_oracle(g, uo, vo, wo)):
            assert np.array_equal(got, want)
    finally:
        reg.close()


def test_four_staggered_normals_from_one_pass_are_exact():
    """calrn_ computes the four staggerings the timestep asks for in one pass
    and parks three of them; the four requests must still return exactly the
    oracle's values, also when the inputs change between two groups."""
    g = Grid(32, 16, 16, 0.1, 0.05, 0.1)
    r = np.random.default_rng(42)
    reg = Registered()
    try:
        px, py, pz = (reg.add(r.standard_normal(g.shape)) for _ in range(3))
        outs = [reg.add(g.zeros(3)) for _ in range(4)]
        dims = [I(g.ni), I(g.nj), I(g.nk)]
        inv = [D(1 / g.dx), D(1 / g.dy), D(1 / g.dz)]
        stag = [(1, 1, 1), (0, 0, 1), (0, 1, 0), (1, 0, 0)]
        x0 = [a.copy() for a in (px, py, pz)]
        for rnd in range(2):
            for (ip, jp, kp), o in zip(stag, outs):
                H.call_gpu("calrn", *dims, *inv, I(ip), I(jp), I(kp), px, py, pz, o)
            for o in outs:
                reg.download(o)
            for (ip, jp, kp), o in zip(stag, outs):
                want = g.zeros(3)
                H.call_oracle("calrn", *dims, *inv, I(ip), I(jp), I(kp), *[a.copy() for a in x0],
                              want)
                assert np.array_equal(o, want), (rnd, ip, jp, kp)
            # new inputs (a host upload): the next group must be recomputed
            x0 = [r.standard_normal(g.shape) for _ in range(3)]
            for a, b in zip((px, py, pz), x0):
                a[...] = b
                H.gpu().felgpu_upload_(ctypes.c_void_p(a.ctypes.data))
        # out-of-order and repeated requests
        for (ip, jp, kp), o in reversed(list(zip(stag, outs))):
            H.call_gpu("calrn", *dims, *inv, I(ip), I(jp), I(kp), px, py, pz, o)
        H.call_gpu("calrn", *dims, *inv, I(0), I(1), I(0), px, py, pz, outs[0])
        for o in outs:
            reg.download(o)
        for (ip, jp, kp), o in zip([(0, 1, 0)] + stag[1:], outs):
            want = g.zeros(3)
            H.call_oracle("calrn", *dims, *inv, I(ip), I(jp), I(kp), *[a.copy() for a in x0], want)
            assert np.array_equal(o, want), ("reordered", ip, jp, kp)
    finally:
        reg.close()


STAG = [(1, 1, 1), (0, 0, 1), (0, 1, 0), (1, 0, 0)]


def csf_chain(call, g, dyinv, q, qn, a):
    """The surface-tension sequence of one bubble (srcs/main.f90:543-571):
    normals, curvature and force of q and of qn, summed into sum / sumn."""
    dims = [I(g.ni), I(g.nj), I(g.nk)]
    inv = [D(1 / g.dx), D(1 / g.dy), D(1 / g.dz)]
    grad = [I(0), I(g.nj)] + dims + [D(1 / g.dx), dyinv, D(1 / g.dz)]
    call("init_q", *dims, *a["sum"])
    call("init_q", *dims, *a["sumn"])
    call("cal_grad_p2a", *grad, q, *a["px"])
    for (ip, jp, kp), o in zip(STAG, a["rn"]):
        call("calrn", *dims, *inv, I(ip), I(jp), I(kp), *a["px"], o)
    call("cal_grad_p2a", *grad, qn, *a["px"])
    for (ip, jp, kp), o in zip(STAG, a["rnn"]):
        call("calrn", *dims, *inv, I(ip), I(jp), I(kp), *a["px"], o)
    call("calrkap", *dims, *inv, *a["rn"], a["rkap"])
    call("calrkap", *dims, *inv, *a["rnn"], a["rkapn"])
    st = [D(1.0), D(0.001), D(0.07)]
    call("cal_grad_p2uvw", *grad, q, *a["px"])
    call("calfst", *dims, *st, *a["px"], a["rkap"], *a["fst"])
    call("cal_grad_p2uvw", *grad, qn, *a["px"])
    call("calfst", *dims, *st, *a["px"], a["rkapn"], *a["fstn"])
    call("sum_fst", *dims, *a["fst"], *a["sum"])
    call("sum_fst", *dims, *a["fstn"], *a["sumn"])


def chain_arrays(g, make):
    return {"px": [make(None) for _ in range(3)], "rn": [make(3) for _ in range(4)],
            "rnn": [make(3) for _ in range(4)], "rkap": make(3), "rkapn": make(3),
            "fst": [make(None) for _ in range(3)], "fstn": [make(None) for _ in range(3)],
            "sum": [make(None) for _ in range(3)], "sumn": [make(None) for _ in range(3)]}


def flat(a):
    out = []
    for k in sorted(a):
        out += a[k] if isinstance(a[k], list) else [a[k]]
    return out


def test_surface_tension_chain_of_the_copied_phase_is_reused_exactly():
    """Step n+1 repeats step n's chain of phin on phi = copy of phin (cpy_ of
    slice 1): the normals, the curvature and the force are taken over from
    the arrays of step n, and the two gradients from the buffers parked when
    the gradient arrays were overwritten (8 calls answered without a
    kernel), and every array equals the oracle's chain bit for bit."""
    g = Grid(24, 12, 16, 0.1, 0.05, 0.1)
    r = np.random.default_rng(43)
    reg = Registered()
    try:
        phi = reg.add(np.stack([H.smooth_phi(g, r), H.smooth_phi(g, r)]))
        phin = reg.add(np.stack([H.smooth_phi(g, r), H.smooth_phi(g, r)]))
        nxt = reg.add(H.smooth_phi(g, r))
        a = chain_arrays(g, lambda m: reg.add(g.zeros(m)))
        dyinv = np.full(g.nj + 6, 1 / g.dy)
        dims = [I(g.ni), I(g.nj), I(g.nk)]
        host_phi, host_phin, host_nxt = phi[1].copy(), phin[1].copy(), nxt.copy()
        csf_chain(H.call_gpu, g, dyinv, phi[1], phin[1], a)
        # the step's end: phi(:,:,:,1) <- phin(:,:,:,1); a new phin
        H.call_gpu("cpy", *dims, phin[1], phi[1])
        H.call_gpu("cpy", *dims, nxt, phin[1])
        H.gpu().felgpu_result_reuse_hits_.restype = ctypes.c_int64
        hits0 = H.gpu().felgpu_result_reuse_hits_()
        csf_chain(H.call_gpu, g, dyinv, phi[1], phin[1], a)
        hits = H.gpu().felgpu_result_reuse_hits_() - hits0
        for x in flat(a):
            reg.download(x)
        want = chain_arrays(g, lambda m: g.zeros(m))
        csf_chain(H.call_oracle, g, dyinv, host_phi, host_phin, want)
        csf_chain(H.call_oracle, g, dyinv, host_phin, host_nxt, want)
        for k in sorted(a):
            for m, (x, y) in enumerate(zip(a[k] if isinstance(a[k], list) else [a[k]],
                                           want[k] if isinstance(want[k], list) else [want[k]])):
                assert same_bits(x, y), (k, m)
        assert hits == 8, hits
    finally:
        reg.close()


RUN_FIELDS = ["u", "v", "w", "p", "uo", "vo", "wo", "po", "phi0", "phi1", "phin1", "rho", "rmu",
              "rhon", "rmun", "src_u", "src_v", "src_w", "adv_u", "adv_uo", "vis_u", "fst_u",
              "fst_un", "fst_w", "sum_fst_u", "sum_fst_vn", "sum_fst_w", "phix", "phiz", "rn_pp",
              "rn_13", "rnn_23", "rkap", "rkapn", "dp", "div"]


@pytest.mark.parametrize("precision,tangent", [(64, False), (32, False), (64, True)],
                         ids=["fp64", "fp32", "tangent"])
def test_runs_with_and_without_reuse_are_bitwise_equal(precision, tangent):
    """Every array of the run after several steps is the same with result
    reuse on (parked results, the memo of the surface-tension chain) as with
    every result recomputed, and reuse is actually exercised."""
    from felmpf import config as C
    from felmpf import sensitivity as S
    from felmpf.driver import GpuRun
    cfg = C.droplet_impact().scaled(32, 16, 32)
    got = {}
    for reuse in (False, True):
        run = GpuRun(cfg, precision=precision, tangent=tangent, reuse=reuse)
        try:
            if tangent:  # derivative along the bottom wall's contact angle
                for name, dirn in S.parameter_seeds(cfg, slip=False)[0].items():
                    run.seed(name, dirn)
            h0 = run.gpu.felgpu_result_reuse_hits_()
            run.steps(4)
            hits = run.gpu.felgpu_result_reuse_hits_() - h0
            got[reuse] = {f: run.field(f) for f in RUN_FIELDS}
            if tangent:
                got[reuse].update({"d_" + f: run.tangent(f) for f in ("u", "w", "phi1")})
        finally:
            run.close()
        assert (hits > 0) == reuse, (reuse, hits)
    for f in got[False]:
        assert same_bits(got[False][f], got[True][f]), f


def test_deferred_zero_and_force_sums_are_exact():
    """init_q_ on device-resident arrays defers the fill, and sum_fst_ onto
    such zeros defers 0.0 + force (srcs/sum_fst.f:28-30); cal_srcu_ reads
    the forces in their place.  Every observable result equals the oracle's
    sequence bit for bit, including -0.0 forces (0.0 + -0.0 = +0.0), a
    force overwritten after the deferral, and an accumulation onto a
    deferred sum."""
    g = Grid(16, 12, 10, 0.1, 0.05, 0.1)
    r = np.random.default_rng(44)
    reg = Registered()
    try:
        base = [r.standard_normal(g.shape) for _ in range(3)] + \
            [r.uniform(0.5, 2.0, g.shape) for _ in range(2)] + \
            [r.standard_normal(g.shape) for _ in range(12)]
        fst = [r.standard_normal(g.shape) for _ in range(6)]
        for f in fst:
            f[r.random(g.shape) < 0.05] = -0.0
        host_fst = [f.copy() for f in fst]
        dev_base = [reg.add(a.copy()) for a in base]
        dev_fst = [reg.add(f) for f in fst]
        dev_sum = [reg.add(r.standard_normal(g.shape)) for _ in range(6)]
        dev_src = [reg.add(g.zeros()) for _ in range(3)]
        dims = [I(g.ni), I(g.nj), I(g.nk)]
        head = [NID.copy(), *dims, D(0.9), D(0.003), D(0.1), D(-0.2)]
        H.call_gpu("init_q", *dims, *dev_sum[:3])
        H.call_gpu("init_q", *dims, *dev_sum[3:])
        H.call_gpu("sum_fst", *dims, *dev_fst[:3], *dev_sum[:3])
        H.call_gpu("sum_fst", *dims, *dev_fst[3:], *dev_sum[3:])
        H.call_gpu("cal_srcu", *head, *dev_base, *dev_sum, *dev_src)
        want_sum = [g.zeros() for _ in range(6)]
        H.call_oracle("sum_fst", *dims, *[f.copy() for f in host_fst[:3]], *want_sum[:3])
        H.call_oracle("sum_fst", *dims, *[f.copy() for f in host_fst[3:]], *want_sum[3:])
        want_src = [g.zeros() for _ in range(3)]
        H.call_oracle("cal_srcu", *head, *[a.copy() for a in base], *want_sum, *want_src)
        for x, y in zip(dev_src, want_src):
            assert same_bits(reg.download(x), y)
        assert not np.any(np.signbit(want_sum[0]) & (want_sum[0] == 0.0))  # -0.0 became +0.0
        # a force rewritten while its sum is deferred: the sum keeps 0.0 + the old force
        H.call_gpu("cpy", *dims, dev_base[5], dev_fst[0])
        assert same_bits(reg.download(dev_sum[0]), want_sum[0])
        # a second accumulation onto a deferred sum (two bubbles)
        H.call_gpu("sum_fst", *dims, *dev_fst[3:], *dev_sum[3:])
        H.call_oracle("sum_fst", *dims, *[f.copy() for f in host_fst[3:]], *want_sum[3:])
        for x, y in zip(dev_sum[1:], want_sum[1:]):
            assert same_bits(reg.download(x), y)
        # a deferred zero that is read
        H.call_gpu("init_q", *dims, *dev_sum[:3])
        assert not np.any(reg.download(dev_sum[1]))
    finally:
        reg.close()


@pytest.mark.parametrize("per_slice", [False, True])
def test_summation_over_registered_slices(per_slice):
    """summation_ (srcs/summation.f:14-50, nbub = 2) on a 3-slice array
    registered as one range or slice by slice, after a whole-slice cpy_ into
    slice 2 (a deferred buffer hand-over when slices are ranges of their
    own): bitwise the oracle's result on the same values."""
    g = Grid(12, 6, 10, 0.1, 0.1, 0.1)
    rng = np.random.default_rng(71)
    q = rng.uniform(-0.2, 0.7, (3,) + g.shape)
    extra = rng.uniform(-0.2, 0.7, g.shape)
    want = q.copy()
    want[2] = extra
    H.call_oracle("summation", I(g.ni), I(g.nj), I(g.nk), want, I(2))
    reg = Registered()
    try:
        if per_slice:
            for m in range(3):
                reg.add(q[m])
        else:
            reg.add(q)
        reg.add(extra)
        H.call_gpu("cpy", I(g.ni), I(g.nj), I(g.nk), extra, q[2])
        H.call_gpu("summation", I(g.ni), I(g.nj), I(g.nk), q, I(2))
        for m in range(3) if per_slice else [None]:
            reg.download(q[m] if per_slice else q)
        assert same_bits(q, want)
    finally:
        reg.close()


@pytest.mark.parametrize("rewrite", [False, True])
def test_pair_dense_sources_equal_the_arrays(rewrite):
    """cal_srcu_ into registered source arrays also stores the sources pair by
    pair, and solu_sor4_ reads that copy while the arrays keep its content
    numbers: the solve equals the one on unregistered arrays bit for bit.
    With rewrite, the host changes src_u after cal_srcu_ (felgpu_upload_):
    the copy is then stale and the solve must read the new values."""
    g = Grid(16, 8, 12, 1 / 16, 1 / 8, 1 / 12)
    r = np.random.default_rng(81)
    ins = [r.standard_normal(g.shape) for _ in range(3)] + \
        [r.uniform(0.5, 2.0, g.shape) for _ in range(2)] + \
        [0.1 * r.standard_normal(g.shape) for _ in range(18)]
    head = [NID.copy(), I(g.ni), I(g.nj), I(g.nk), D(0.9), D(0.003), D(0.1), D(-0.2)]
    rmu = r.uniform(0.01, 0.2, g.shape)
    coef = [g.zeros() for _ in range(45)]
    H.call_oracle("cal_arith_coef_vis", I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy),
                  D(1 / g.dz), rmu, *coef)
    uvw0 = [H.smooth_vel(g, r) for _ in range(3)]
    phi = r.uniform(0, 1, g.shape)
    l1 = np.full(g.shape, 0.05)

    def solve(src):
        uvw = [x.copy() for x in uvw0]
        H.call_gpu("solu_sor4", I(0), I(0), NID.copy(), I(1), I(g.ni), I(g.nj), I(g.nk),
                   np.zeros(8, np.int32), np.zeros(8), np.zeros(8), I(1), I(0), D(0.003),
                   ins[3].copy(), ins[4].copy(), *[c.copy() for c in coef], *src, *uvw,
                   D(0.25), D(g.dy), l1.copy(), l1.copy(), phi.copy())
        return uvw

    # unregistered sources: the solve reads the arrays themselves
    plain = [g.zeros() for _ in range(3)]
    H.call_gpu("cal_srcu", *head, *[x.copy() for x in ins], *plain)
    if rewrite:
        plain[0] *= 0.5
    want = solve(plain)
    reg = Registered()
    try:
        src = [reg.add(g.zeros()) for _ in range(3)]
        H.call_gpu("cal_srcu", *head, *[x.copy() for x in ins], *src)
        if rewrite:
            reg.download(src[0])
            src[0] *= 0.5
            H.gpu().felgpu_upload_(ctypes.c_void_p(src[0].ctypes.data))
        got = solve(src)
        for a, b in zip(got, want):
            assert same_bits(a, b)
    finally:
        reg.close()


@pytest.mark.parametrize("dims", [(130, 16, 40), (12, 10, 20)])
def test_normals_and_curvature_in_one_pass(dims):
    """The first calrn_ request on registered gradients forms all four
    staggered normals and their curvature (srcs/calrn.f, srcs/calrkap.f) in
    one pass; calrkap_ on the four delivered normals takes that curvature.
    Normals and curvature equal the separate per-request kernels on
    unregistered arrays bit for bit; several x, y and k tiles are crossed.
    After one input is rewritten, calrkap_ computes from the arrays."""
    ni, nj, nk = dims
    g = Grid(ni, nj, nk, 1 / ni, 1 / nj, 1 / nk)
    rng = np.random.default_rng(91)
    phi = H.smooth_phi(g, rng)
    grads = [g.zeros() for _ in range(3)]
    ID, sv = I(0), I(g.nj)
    dyinv = np.full(g.nj + 6, 1 / g.dy)
    H.call_oracle("cal_grad_p2a", ID, sv, I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), dyinv,
                  D(1 / g.dz), phi, *grads)
    stag = [(1, 1, 1), (0, 0, 1), (0, 1, 0), (1, 0, 0)]
    sc = (D(1 / g.dx), D(1 / g.dy), D(1 / g.dz))
    dims_ = (I(g.ni), I(g.nj), I(g.nk))

    lib = H.gpu()
    lib.felgpu_stash_hits_.restype = ctypes.c_int64

    def run(registered, rewrite=False):
        reg = Registered()
        h0 = lib.felgpu_stash_hits_()
        try:
            gr = [reg.add(x.copy()) if registered else x.copy() for x in grads]
            rn = [np.zeros((3,) + g.shape) for _ in range(4)]
            kap = np.zeros((3,) + g.shape)
            if registered:
                rn = [reg.add(x) for x in rn]
                kap = reg.add(kap)
            for (ip, jp, kp), out in zip(stag, rn):
                H.call_gpu("calrn", *dims_, *sc, I(ip), I(jp), I(kp), *gr, out)
            if rewrite:
                reg.download(rn[2])
                rn[2][1, 3, 4, 5] += 1.0
                H.gpu().felgpu_upload_(ctypes.c_void_p(rn[2].ctypes.data))
            H.call_gpu("calrkap", *dims_, *sc, *rn, kap)
            if registered:
                for x in rn:
                    reg.download(x)
                reg.download(kap)
            # three normals and (unless an input was rewritten) the curvature
            # came from the first request's pass
            want_hits = (3 + (0 if rewrite else 1)) if registered else 0
            assert lib.felgpu_stash_hits_() - h0 == want_hits
            return [x.copy() for x in rn], kap.copy()
        finally:
            reg.close()

    rn_a, kap_a = run(False)
    rn_b, kap_b = run(True)
    for a, b in zip(rn_a, rn_b):
        assert same_bits(a, b)
    assert same_bits(kap_a, kap_b)
    assert np.abs(kap_a).max() > 0
    # a rewritten input: the curvature follows the arrays
    rn_c, kap_c = run(True, rewrite=True)
    rn_a[2][1, 3, 4, 5] += 1.0
    want = np.zeros((3,) + g.shape)
    H.call_gpu("calrkap", *dims_, *sc, *rn_a, want)
    assert same_bits(kap_c, want)
