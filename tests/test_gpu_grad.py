"""Differentiable registration on the GPU (pyflowreg_amd.autograd): each HIP
forward/backward pair against the pure-torch autograd restatement
(tests/grad_reference.py), the single-level forward against the oracle, and
the whole-pyramid gradient against central FP64 finite differences.

Tolerances: stage forwards agree with torch to 1e-11 relative (the same
float64 formulas in another evaluation order); gradients of one level solve
to 1e-8 and of a whole pyramid to 1e-5 relative (rounding through 50 sweeps
per level and sums in another order); central differences to 1e-3
relative, the stated target for the adjoint gradient, at EVERY step of a
decreasing sequence spanning a factor 10 (per direction, inside the window
where the differences have converged: above it kinks of the map, below it
float64 rounding of the loss dominate), so that one lucky step cannot pass a
wrong adjoint.  h = 1e-4 is not usable: the
registration map is piecewise smooth (medians, border crossings, the stiff
robust weight) with kinks far closer than 1e-4 in these directions, so
differences at that step do not estimate a derivative of this map
(whatever computes the gradient); the pyramid check therefore runs at
a_data = 0.7, and the default a_data = 0.45 is checked by finite
differences on one level and against torch autograd on the full pyramid.
"""

import numpy as np
import pytest
import torch

import grad_reference as R

pytestmark = pytest.mark.gpu

F64 = torch.float64


def _rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-300))


def _frames(T, H, W, C, seed):
    """Preprocessed synthetic frames, float32-valued float64: fixed (H, W, C)
    and moving (T, H, W, C) numpy arrays."""
    from oracle import oracle as O
    from pyflowreg_amd import synth
    vid, _ = synth.video(T + 1, H, W, C, seed=seed, max_disp=2.0)
    raw = vid.double().numpy()
    fixed = O.preprocess(raw[0]).astype(np.float32).astype(np.float64)
    moving = O.preprocess(raw[1:], raw[0]).astype(np.float32).astype(np.float64)
    return fixed, moving


def _planes(fixed, moving):
    f = torch.from_numpy(fixed).permute(2, 0, 1).contiguous()
    m = torch.from_numpy(moving).permute(0, 3, 1, 2).contiguous()
    return f, m


def _cuda(*ts):
    return [t.to("cuda").contiguous() for t in ts]


def _skew_index(h, w):
    """Slot of cell (r, c) in the solver's skewed plane: (r + c) * hs + r."""
    hs = 64 * ((h + 63) // 64)
    r = torch.arange(h)[:, None]
    c = torch.arange(w)[None, :]
    return ((r + c) * hs + r).reshape(-1), (h + w - 1) * hs


def _sk(x):
    """(..., K, h, w) -> (..., plane, K): each cell's K values interleaved in
    its slot, zero padding slots."""
    K, h, w = x.shape[-3:]
    idx, P = _skew_index(h, w)
    out = torch.zeros(x.shape[:-3] + (P, K), dtype=x.dtype, device=x.device)
    out[..., idx.to(x.device), :] = x.reshape(x.shape[:-3] + (K, h * w)).transpose(-1, -2)
    return out


def _unsk(x, h, w):
    """(..., plane, K) -> (..., K, h, w)."""
    idx, _ = _skew_index(h, w)
    return x[..., idx.to(x.device), :].transpose(-1, -2).reshape(x.shape[:-2] + (x.shape[-1], h, w))


# ------------------------------------------------------------------ stages
def test_motion_tensor_forward_and_backward(gpu):
    from pyflowreg_amd import _lib
    fixed, moving = _frames(2, 20, 26, 2, seed=21)
    f, m = _planes(fixed, moving)
    T, C, h, w = m.shape
    hy, hx = 1.25, 1.1
    mt = m.clone().requires_grad_(True)
    Jr = R.motion_tensor_gc(f, mt, hy, hx)
    gJ = torch.randn(Jr.shape, dtype=F64, generator=torch.Generator().manual_seed(3))
    (Jr * gJ).sum().backward()
    fd, md, gJd = _cuda(f, m, gJ)
    J = torch.empty((T, 6, C, h, w), dtype=F64, device="cuda")
    g_m = torch.zeros_like(md)
    L = _lib.lib()
    _lib.check(L.frh_grad_tensor(fd.data_ptr(), md.data_ptr(), T, C, h, w, hy, hx, J.data_ptr(),
                                 gJd.data_ptr(), g_m.data_ptr(), _lib.stream_handle()), "tensor")
    torch.cuda.synchronize()
    assert _rel(J, Jr) < 1e-12
    assert _rel(g_m, mt.grad) < 1e-10


@pytest.mark.parametrize("psi_one", [False, True])
def test_coefficients_forward_and_backward(gpu, psi_one):
    from pyflowreg_amd import _lib
    g = torch.Generator().manual_seed(5)
    fixed, moving = _frames(2, 18, 23, 2, seed=22)
    f, m = _planes(fixed, moving)
    T, C, h, w = m.shape
    J = R.motion_tensor_gc(f, m, 1.0, 1.0).detach()
    W = torch.rand((C, h, w), dtype=F64, generator=g) * 0.5 + 0.25
    x = (torch.randn((T, 2, h, w), dtype=F64, generator=g) * 0.3).requires_grad_(True)
    u0 = (torch.randn((T, 2, h, w), dtype=F64, generator=g) * 0.5).requires_grad_(True)
    Jt = J.clone().requires_grad_(True)
    ax = torch.tensor(1.7, dtype=F64, requires_grad=True)
    ay = torch.tensor(2.3, dtype=F64, requires_grad=True)
    a = torch.tensor([0.45, 0.6], dtype=F64, requires_grad=True)
    K = R.coefficients(x, u0, Jt, W, ax, ay, a, psi_one)
    gK = torch.randn(K.shape, dtype=F64, generator=g)
    (K * gK).sum().backward()

    xd, u0d, Jd, Wd, ad, gKd = _cuda(_sk(x.detach()), u0.detach(), J, W, a.detach(), _sk(gK))
    _, P = _skew_index(h, w)
    coef = torch.empty((T, P, 6), dtype=F64, device="cuda")
    g_x = torch.zeros_like(xd)
    g_u0 = torch.zeros_like(u0d)
    g_J = torch.zeros_like(Jd)
    g_par = torch.zeros((T, 2 + C, P), dtype=F64, device="cuda")
    L = _lib.lib()
    _lib.check(L.frh_grad_coef(xd.data_ptr(), u0d.data_ptr(), Jd.data_ptr(), Wd.data_ptr(), T, C, h, w,
                               1.7, 2.3, ad.data_ptr(), int(psi_one), coef.data_ptr(), gKd.data_ptr(),
                               g_x.data_ptr(), g_u0.data_ptr(), g_J.data_ptr(), g_par.data_ptr(),
                               _lib.stream_handle()), "coef")
    torch.cuda.synchronize()
    assert _rel(_unsk(coef, h, w), K) < 1e-12
    if psi_one:
        assert float(g_x.abs().max()) == 0.0
    else:
        assert _rel(_unsk(g_x, h, w), x.grad) < 1e-10
    assert _rel(g_u0, u0.grad) < 1e-12
    assert _rel(g_J, Jt.grad) < 1e-10
    assert abs(float(g_par[:, 0].sum()) - float(ax.grad)) <= 1e-10 * abs(float(ax.grad))
    assert abs(float(g_par[:, 1].sum()) - float(ay.grad)) <= 1e-10 * abs(float(ay.grad))
    ga = g_par[:, 2:].sum(dim=(0, 2)).cpu()
    if psi_one:
        assert float(ga.abs().max()) == 0.0
    else:
        assert _rel(ga, a.grad) < 1e-10


def test_sweep_forward_and_adjoint(gpu):
    from pyflowreg_amd import _lib
    g = torch.Generator().manual_seed(7)
    fixed, moving = _frames(2, 17, 21, 1, seed=23)
    f, m = _planes(fixed, moving)
    T, C, h, w = m.shape
    J = R.motion_tensor_gc(f, m, 1.0, 1.0)
    W = torch.full((C, h, w), 1.0, dtype=F64)
    x0 = torch.randn((T, 2, h, w), dtype=F64, generator=g) * 0.3
    u0 = torch.randn((T, 2, h, w), dtype=F64, generator=g) * 0.3
    coef = R.coefficients(x0, u0, J, W, 2.0, 1.5, torch.tensor([0.45], dtype=F64), False).detach()
    coef_t = coef.clone().requires_grad_(True)
    x = x0.clone().requires_grad_(True)
    ax = torch.tensor(2.0, dtype=F64, requires_grad=True)
    ay = torch.tensor(1.5, dtype=F64, requires_grad=True)
    y = R.sor_sweep(coef_t, x, ax, ay)
    lam = torch.randn(y.shape, dtype=F64, generator=g)
    (y * lam).sum().backward()

    cd, xd, lamd = _cuda(_sk(coef), _sk(x0), _sk(lam))
    _, P = _skew_index(h, w)
    out = torch.empty_like(xd)
    lam_in = torch.empty_like(xd)
    g_coef = torch.zeros_like(cd)
    g_axy = torch.zeros((T, 2, P), dtype=F64, device="cuda")
    L = _lib.lib()
    st = _lib.stream_handle()
    _lib.check(L.frh_grad_sweep(cd.data_ptr(), xd.data_ptr(), out.data_ptr(), T, h, w, 2.0, 1.5, st), "sweep")
    _lib.check(L.frh_grad_sweep_adjoint(cd.data_ptr(), xd.data_ptr(), out.data_ptr(), lamd.data_ptr(),
                                        lam_in.data_ptr(), g_coef.data_ptr(), 0, g_axy.data_ptr(),
                                        T, h, w, 2.0, 1.5, st), "adjoint")
    # g_coef_set: the coefficient adjoints are stored, whatever the buffer held
    g_set = torch.full_like(cd, 7.0)
    g_axy2 = torch.zeros_like(g_axy)
    _lib.check(L.frh_grad_sweep_adjoint(cd.data_ptr(), xd.data_ptr(), out.data_ptr(), lamd.data_ptr(),
                                        lam_in.data_ptr(), g_set.data_ptr(), 1, g_axy2.data_ptr(),
                                        T, h, w, 2.0, 1.5, st), "adjoint (set)")
    torch.cuda.synchronize()
    assert torch.equal(_unsk(g_set, h, w), _unsk(g_coef, h, w))
    assert torch.equal(g_axy2, g_axy)
    assert _rel(_unsk(out, h, w), y) < 1e-12
    assert _rel(_unsk(lam_in, h, w), x.grad) < 1e-11
    assert _rel(_unsk(g_coef, h, w), coef_t.grad) < 1e-11
    assert abs(float(g_axy[:, 0].sum()) - float(ax.grad)) <= 1e-10 * abs(float(ax.grad))
    assert abs(float(g_axy[:, 1].sum()) - float(ay.grad)) <= 1e-10 * abs(float(ay.grad))


def test_sweep_one_row_and_one_column(gpu):
    """Degenerate levels: a single row or column (every cell on the level
    border), against the torch sweep."""
    from pyflowreg_amd import _lib
    g = torch.Generator().manual_seed(8)
    L = _lib.lib()
    for h, w in ((1, 9), (9, 1), (1, 1), (70, 3), (130, 5)):
        coef = torch.rand((2, 6, h, w), dtype=F64, generator=g) * 0.5
        x = torch.randn((2, 2, h, w), dtype=F64, generator=g)
        y = R.sor_sweep(coef, x, 1.3, 0.7)
        cd, xd = _cuda(_sk(coef), _sk(x))
        out = torch.empty_like(xd)
        _lib.check(L.frh_grad_sweep(cd.data_ptr(), xd.data_ptr(), out.data_ptr(), 2, h, w, 1.3, 0.7,
                                    _lib.stream_handle()), "sweep")
        torch.cuda.synchronize()
        assert _rel(_unsk(out, h, w), y) < 1e-13, (h, w)


def test_sweep_adjoint_several_waves(gpu):
    """Levels of 65..300 rows put a level's rows into 2..5 waves, whose edge
    lanes exchange through LDS in both directions."""
    from pyflowreg_amd import _lib
    g = torch.Generator().manual_seed(15)
    L = _lib.lib()
    for h, w in ((65, 7), (128, 4), (300, 6)):
        coef = torch.rand((1, 6, h, w), dtype=F64, generator=g) * 0.5
        x = torch.randn((1, 2, h, w), dtype=F64, generator=g).requires_grad_(True)
        ct = coef.clone().requires_grad_(True)
        y = R.sor_sweep(ct, x, 1.1, 0.9)
        lam = torch.randn(y.shape, dtype=F64, generator=g)
        (y * lam).sum().backward()
        _, P = _skew_index(h, w)
        cd, xd, lamd = _cuda(_sk(coef), _sk(x.detach()), _sk(lam))
        out, lam_in = torch.empty_like(xd), torch.empty_like(xd)
        g_coef = torch.zeros_like(cd)
        g_axy = torch.zeros((1, 2, P), dtype=F64, device="cuda")
        st = _lib.stream_handle()
        _lib.check(L.frh_grad_sweep(cd.data_ptr(), xd.data_ptr(), out.data_ptr(), 1, h, w, 1.1, 0.9, st), "sweep")
        _lib.check(L.frh_grad_sweep_adjoint(cd.data_ptr(), xd.data_ptr(), out.data_ptr(), lamd.data_ptr(),
                                            lam_in.data_ptr(), g_coef.data_ptr(), 1, g_axy.data_ptr(),
                                            1, h, w, 1.1, 0.9, st), "adjoint")
        torch.cuda.synchronize()
        assert _rel(_unsk(out, h, w), y) < 1e-12, (h, w)
        assert _rel(_unsk(lam_in, h, w), x.grad) < 1e-11, (h, w)
        assert _rel(_unsk(g_coef, h, w), ct.grad) < 1e-11, (h, w)


def _sweeps_case(L, _lib, g, T, h, w, ax, ay):
    coef = torch.rand((T, 6, h, w), dtype=F64, generator=g) * 0.5
    x = torch.randn((T, 2, h, w), dtype=F64, generator=g)
    cd, xd = _cuda(_sk(coef), _sk(x))
    return cd, xd


def test_pipelined_sweeps_equal_single_sweeps(gpu):
    """frh_grad_sweeps against the same number of single-sweep launches,
    bit-exact, for every state when write_all, for the last when not (also
    in place), at 1..9 sweeps and levels of 1..700 rows (the 1024-lane form
    at 700): without scratch (up to 5 sweeps per wavefront pass, LDS exchange
    of every sweep's edge values between waves) and with scratch (a chain of
    work-groups per frame under progress flags; frame counts not multiples
    of 8, and 33 sweeps: two chain launches)."""
    from pyflowreg_amd import _lib
    g = torch.Generator().manual_seed(31)
    L = _lib.lib()
    st = _lib.stream_handle()
    ax, ay = 1.2, 0.8
    for (h, w, T, ns) in ((17, 21, 2, (1, 2, 3, 4, 5, 9)), (1, 9, 2, (2, 5)), (9, 1, 1, (2, 5)), (1, 1, 1, (3,)),
                          (65, 7, 3, (2, 5, 9)), (130, 5, 2, (4, 5)), (300, 40, 1, (5,)), (512, 37, 9, (3, 5)),
                          (700, 12, 1, (2, 3)), (19, 6, 2, (33,))):
        cd, xd = _sweeps_case(L, _lib, g, T, h, w, ax, ay)
        for n in ns:
            ref = torch.empty((n,) + xd.shape, dtype=F64, device="cuda")
            src = xd
            for j in range(n):
                _lib.check(L.frh_grad_sweep(cd.data_ptr(), src.data_ptr(), ref[j].data_ptr(), T, h, w, ax, ay,
                                            st), "sweep")
                src = ref[j]
            nb = L.frh_grad_sweeps_scratch_bytes(n, T)
            scratch = torch.zeros(nb, dtype=torch.uint8, device="cuda")
            for sp, sb in ((None, 0), (scratch.data_ptr(), nb)):
                case = (h, w, T, n, sp is not None)
                allx = torch.full_like(ref, float("nan"))
                _lib.check(L.frh_grad_sweeps(cd.data_ptr(), xd.data_ptr(), allx.data_ptr(), xd.numel(), n, 1, T, h,
                                             w, ax, ay, sp, sb, st), "sweeps (all)")
                last = torch.full_like(xd, float("nan"))
                _lib.check(L.frh_grad_sweeps(cd.data_ptr(), xd.data_ptr(), last.data_ptr(), 0, n, 0, T, h, w,
                                             ax, ay, sp, sb, st), "sweeps (last)")
                inplace = xd.clone()
                _lib.check(L.frh_grad_sweeps(cd.data_ptr(), inplace.data_ptr(), inplace.data_ptr(), 0, n, 0, T, h,
                                             w, ax, ay, sp, sb, st), "sweeps (in place)")
                torch.cuda.synchronize()
                # (padding slots are never written)
                assert torch.equal(_unsk(allx, h, w), _unsk(ref, h, w)), case
                assert torch.equal(_unsk(last, h, w), _unsk(ref[-1], h, w)), case
                assert torch.equal(_unsk(inplace, h, w), _unsk(ref[-1], h, w)), case
            assert int(scratch[:4].view(torch.int32)[0]) == 0, (h, w, T, n)
    # against the torch sweep, 5 sweeps in one pass
    coef = torch.rand((1, 6, 70, 33), dtype=F64, generator=g) * 0.5
    x = torch.randn((1, 2, 70, 33), dtype=F64, generator=g)
    y = x
    for _ in range(5):
        y = R.sor_sweep(coef, y, ax, ay)
    cd, xd = _cuda(_sk(coef), _sk(x))
    out = torch.empty_like(xd)
    _lib.check(L.frh_grad_sweeps(cd.data_ptr(), xd.data_ptr(), out.data_ptr(), 0, 5, 0, 1, 70, 33, ax, ay, None, 0,
                                 st), "sweeps")
    torch.cuda.synchronize()
    assert _rel(_unsk(out, 70, 33), y) < 1e-12


def test_chained_adjoint_equals_single_adjoints(gpu):
    """frh_grad_sweeps_adjoint (one work-group per sweep and frame, lam_in
    handed on through global memory under progress flags) against the same
    adjoints as single launches from the block's last sweep to its first:
    bit-exact lam, coefficient adjoints (set and accumulated) and row partials,
    for frame counts that are not multiples of 8, levels of 1..700 rows (the
    1024-lane form at 700) and blocks longer than one launch (33, 65 sweeps:
    hand-over planes between launches); no stage gave up waiting."""
    from pyflowreg_amd import _lib
    g = torch.Generator().manual_seed(41)
    L = _lib.lib()
    st = _lib.stream_handle()
    ax, ay = 1.1, 0.9
    nan = float("nan")
    for (h, w, T, ns) in ((17, 21, 3, (1, 2, 3, 5)), (1, 1, 1, (2, 5)), (1, 9, 2, (3,)), (9, 1, 2, (3,)),
                          (65, 7, 9, (2, 5)), (300, 40, 2, (5,)), (512, 37, 1, (4, 5)), (700, 12, 1, (3,)),
                          (17, 5, 2, (33, 65))):
        coef = torch.rand((T, 6, h, w), dtype=F64, generator=g) * 0.5
        x = torch.randn((T, 2, h, w), dtype=F64, generator=g)
        lam = torch.randn((T, 2, h, w), dtype=F64, generator=g)
        cd, xd, lamd = _cuda(_sk(coef), _sk(x), _sk(lam))
        _, P = _skew_index(h, w)
        for n in ns:
            xs = torch.empty((n,) + xd.shape, dtype=F64, device="cuda")
            src = xd
            for j in range(n):
                _lib.check(L.frh_grad_sweep(cd.data_ptr(), src.data_ptr(), xs[j].data_ptr(), T, h, w, ax, ay, st),
                           "sweep")
                src = xs[j]
            for gset in (0, 1):
                g_ref = torch.full_like(cd, 0.25)
                gaxy_ref = torch.full((T, 2, P), 0.5, dtype=F64, device="cuda")
                lam_a, lam_b = lamd.clone(), torch.empty_like(lamd)
                for j in range(n - 1, -1, -1):
                    x_in = xd if j == 0 else xs[j - 1]
                    _lib.check(L.frh_grad_sweep_adjoint(cd.data_ptr(), x_in.data_ptr(), xs[j].data_ptr(),
                                                        lam_a.data_ptr(), lam_b.data_ptr(), g_ref.data_ptr(),
                                                        int(gset and j == n - 1), gaxy_ref.data_ptr(), T, h, w,
                                                        ax, ay, st), "adjoint")
                    lam_a, lam_b = lam_b, lam_a
                nb = L.frh_grad_sweeps_adjoint_scratch_bytes(n, T, h, w)
                scratch = torch.zeros(nb, dtype=torch.uint8, device="cuda")
                g_ch = torch.full_like(cd, 0.25)
                gaxy_ch = torch.full((T, 2, P), 0.5, dtype=F64, device="cuda")
                lam_out = torch.full_like(lamd, nan)
                _lib.check(L.frh_grad_sweeps_adjoint(cd.data_ptr(), xd.data_ptr(), xs.data_ptr(), xd.numel(), n,
                                                     lamd.data_ptr(), lam_out.data_ptr(), g_ch.data_ptr(), gset,
                                                     gaxy_ch.data_ptr(), T, h, w, ax, ay, scratch.data_ptr(), nb,
                                                     st), "chained adjoint")
                torch.cuda.synchronize()
                case = (h, w, T, n, gset)
                assert int(scratch[:4].view(torch.int32)[0]) == 0, case
                assert torch.equal(_unsk(lam_out, h, w), _unsk(lam_a, h, w)), case
                assert torch.equal(g_ch, g_ref), case
                assert torch.equal(gaxy_ch, gaxy_ref), case


def test_skew_conversion_round_trip(gpu):
    from pyflowreg_amd import autograd as AG
    x = torch.randn((3, 2, 37, 70), dtype=F64, device="cuda")
    s = AG.to_skew(x)
    assert s.shape == (3, AG.skew_plane(37, 70), 2)
    assert torch.equal(s.cpu(), _sk(x.cpu()))
    assert torch.equal(AG.from_skew(s, 37, 70), x)


def test_median_forward_and_backward(gpu):
    from pyflowreg_amd import _lib
    g = torch.Generator().manual_seed(9)
    x = torch.randn((3, 11, 14), dtype=F64, generator=g).requires_grad_(True)
    y = R.median5(x)
    gy = torch.randn(y.shape, dtype=F64, generator=g)
    (y * gy).sum().backward()
    xd, gyd = _cuda(x.detach(), gy)
    out = torch.empty_like(xd)
    gx = torch.zeros_like(xd)
    L = _lib.lib()
    _lib.check(L.frh_grad_median(xd.data_ptr(), 3, 11, 14, out.data_ptr(), gyd.data_ptr(), gx.data_ptr(),
                                 _lib.stream_handle()), "median")
    torch.cuda.synchronize()
    assert torch.equal(out.cpu(), y.detach())
    assert _rel(gx, x.grad) < 1e-14


def test_continuous_warp_forward_and_backward(gpu):
    from pyflowreg_amd import autograd as AG
    g = torch.Generator().manual_seed(10)
    T, C, H, W = 2, 2, 13, 17
    img = torch.rand((T, C, H, W), dtype=F64, generator=g)
    fill = torch.rand((C, H, W), dtype=F64, generator=g)
    # displacements up to 3 px: some samples leave the image, some clip
    u = (torch.rand((T, H, W), dtype=F64, generator=g) - 0.5) * 6
    v = (torch.rand((T, H, W), dtype=F64, generator=g) - 0.5) * 6
    it, ut, vt = (t.clone().requires_grad_(True) for t in (img, u, v))
    ft = fill.clone().requires_grad_(True)
    out_r = R.warp(it, ut, vt, ft)
    go = torch.randn(out_r.shape, dtype=F64, generator=g)
    (out_r * go).sum().backward()
    id_, fd, ud, vd, god = _cuda(img, fill, u, v, go)
    id_.requires_grad_(True)
    fd.requires_grad_(True)
    ud.requires_grad_(True)
    vd.requires_grad_(True)
    out = AG.warp(id_, ud, vd, fd)
    (out * god).sum().backward()
    assert _rel(out, out_r) < 1e-13
    assert _rel(id_.grad, it.grad) < 1e-12
    assert _rel(ud.grad, ut.grad) < 1e-12
    assert _rel(vd.grad, vt.grad) < 1e-12
    assert _rel(fd.grad, ft.grad) < 1e-14


def test_continuous_warp_integer_shift_and_zero_flow(gpu):
    """Zero flow is the identity and an integer shift moves pixels exactly;
    samples that leave the image take the fill (core/warping.py:160-234)."""
    from pyflowreg_amd import autograd as AG
    img = torch.rand((1, 1, 9, 12), dtype=F64, device="cuda")
    fill = torch.rand((1, 9, 12), dtype=F64, device="cuda")
    z = torch.zeros((1, 9, 12), dtype=F64, device="cuda")
    assert torch.equal(AG.warp(img, z, z, fill), img)
    out = AG.warp(img, z + 2.0, z, fill)
    assert torch.equal(out[0, 0, :, :-2], img[0, 0, :, 2:])
    assert torch.equal(out[0, 0, :, -2:], fill[0, :, -2:])


# ------------------------------------------------------------- level solve
def test_level_solve_matches_torch_autograd(gpu):
    from pyflowreg_amd import autograd as AG
    g = torch.Generator().manual_seed(12)
    fixed, moving = _frames(2, 21, 19, 2, seed=24)
    f, m = _planes(fixed, moving)
    T, C, h, w = m.shape
    u0 = torch.randn((T, 2, h, w), dtype=F64, generator=g) * 0.4
    W = torch.full((C, h, w), 0.5, dtype=F64)
    hx, hy = 1.3, 1.2
    leaves = dict(m=m.clone(), u0=u0.clone(), ax=torch.tensor(2.0 / hx ** 2, dtype=F64),
                  ay=torch.tensor(2.5 / hy ** 2, dtype=F64), a=torch.tensor([0.45, 0.55], dtype=F64))
    ref = {k: v.clone().requires_grad_(True) for k, v in leaves.items()}
    out_r = R.level_solve(f, ref["m"], ref["u0"], ref["ax"], ref["ay"], ref["a"], W, 13, 5, hx, hy, True)
    go = torch.randn(out_r.shape, dtype=F64, generator=g)
    (out_r * go).sum().backward()
    gd = {k: v.to("cuda").clone().requires_grad_(True) for k, v in leaves.items()}
    fd, Wd, god = _cuda(f, W, go)
    out = AG.LevelSolve.apply(fd, gd["m"], gd["u0"], gd["ax"], gd["ay"], gd["a"], Wd, 13, 5, hx, hy, True)
    (out * god).sum().backward()
    assert _rel(out, out_r) < 1e-11
    for k in leaves:
        assert _rel(gd[k].grad, ref[k].grad) < 1e-8, k


# --------------------------------------------------------- whole pyramid
def test_single_level_forward_matches_oracle(gpu, oracle):
    """levels=0: no warp and float32-valued inputs, so the differentiable map
    is the reference's get_displacement(levels=0) up to rounding."""
    from pyflowreg_amd import autograd as AG
    fixed, moving = _frames(2, 30, 26, 2, seed=25)
    w = AG.get_displacement(torch.from_numpy(fixed).cuda(), torch.from_numpy(moving).cuda(), levels=0)
    for t in range(2):
        ref = oracle.get_displacement(fixed, moving[t], levels=0)
        assert np.abs(w[t].cpu().numpy() - ref).max() < 1e-10 * max(1.0, np.abs(ref).max())


def test_pyramid_forward_and_gradient_match_torch_autograd(gpu, oracle):
    from pyflowreg_amd import autograd as AG
    fixed, moving = _frames(2, 40, 48, 2, seed=26)
    f, m = _planes(fixed, moving)
    mask = (torch.rand((40, 48), dtype=F64, generator=torch.Generator().manual_seed(4)) > 0.2).double()
    # reference
    alpha_r = torch.tensor([2.0, 2.0], dtype=F64, requires_grad=True)
    a_r = torch.tensor([0.45, 0.45], dtype=F64, requires_grad=True)
    m_r = m.clone().requires_grad_(True)
    flow_r = R.get_displacement(f, m_r, alpha_r, a_r, oracle, iterations=20)
    loss_r = R.loss(f, m_r, flow_r, mask)
    loss_r.backward()
    # GPU
    alpha = torch.tensor([2.0, 2.0], dtype=F64, device="cuda", requires_grad=True)
    a = torch.tensor([0.45, 0.45], dtype=F64, device="cuda", requires_grad=True)
    mv = torch.from_numpy(moving).cuda().requires_grad_(True)
    fx = torch.from_numpy(fixed).cuda()
    flow = AG.get_displacement(fx, mv, alpha=alpha, a_data=a, iterations=20)
    loss = AG.registration_loss(fx, mv, flow, mask.cuda())
    loss.backward()
    errs = dict(flow=_rel(flow.permute(0, 3, 1, 2), flow_r), alpha=_rel(alpha.grad, alpha_r.grad),
                a_data=_rel(a.grad, a_r.grad), moving=_rel(mv.grad.permute(0, 3, 1, 2), m_r.grad))
    print("relative errors vs torch autograd:", errs, flush=True)
    assert errs["flow"] < 1e-10
    assert abs(float(loss.detach()) - float(loss_r.detach())) <= 1e-10 * abs(float(loss_r.detach()))
    # 12 levels of sums in another order (atomics in the warp and tensor
    # backward): gradients agree to ~1e-8
    assert errs["alpha"] < 1e-6 and errs["a_data"] < 1e-6 and errs["moving"] < 1e-6


def _fd_check(fixed, moving, mask, directions, steps, a_data=0.45, **opts):
    """Analytic directional derivatives of the loss and central differences
    at each step in ``steps[name]``: [(name, analytic, {h: fd})]."""
    from pyflowreg_amd import autograd as AG
    fx = torch.from_numpy(fixed).cuda()
    base = dict(alpha=torch.tensor([2.0, 2.0], dtype=F64, device="cuda"),
                a_data=torch.tensor([a_data] * fixed.shape[2], dtype=F64, device="cuda"),
                moving=torch.from_numpy(moving).cuda())
    mk = torch.as_tensor(mask, device="cuda")

    def loss_of(p):
        flow = AG.get_displacement(fx, p["moving"], alpha=p["alpha"], a_data=p["a_data"], **opts)
        return AG.registration_loss(fx, p["moving"], flow, mk)

    leaves = {k: v.clone().requires_grad_(True) for k, v in base.items()}
    loss_of(leaves).backward()
    out = []
    for name, d in directions:
        d = d.to("cuda")
        analytic = float((leaves[name].grad * d).sum())
        fds = {}
        with torch.no_grad():
            for h in steps[name]:
                lp = float(loss_of({**base, name: base[name] + h * d}))
                lm = float(loss_of({**base, name: base[name] - h * d}))
                fds[h] = (lp - lm) / (2 * h)
        out.append((name, analytic, fds))
    return out


def _directions(moving, seed):
    g = torch.Generator().manual_seed(seed)
    dirs = [("alpha", torch.tensor([1.0, 0.0], dtype=F64)), ("alpha", torch.tensor([0.0, 1.0], dtype=F64)),
            ("a_data", torch.tensor([1.0, 0.0], dtype=F64)), ("a_data", torch.tensor([0.0, 1.0], dtype=F64))]
    for _ in range(2):
        dirs.append(("moving", torch.randn(moving.shape, dtype=F64, generator=g)))
    return dirs


def _assert_fd(results, tol=1e-3):
    """Every step of the sequence must agree with the analytic derivative,
    and the differences must agree with each other (a converged sequence)."""
    for name, an, fds in results:
        rel = {h: abs(an - fd) / max(abs(an), 1e-300) for h, fd in fds.items()}
        print(f"{name}: analytic {an:.10e} " + " ".join(f"h={h:g}:{r:.1e}" for h, r in rel.items()), flush=True)
    for name, an, fds in results:
        rel = {h: abs(an - fd) / max(abs(an), 1e-300) for h, fd in fds.items()}
        assert max(rel.values()) <= tol, (name, an, fds)
        vals = list(fds.values())
        spread = (max(vals) - min(vals)) / max(abs(an), 1e-300)
        assert spread <= tol, (name, "finite differences not converged", fds)


@pytest.mark.timeout(600)
def test_adjoint_gradient_vs_central_differences_128(gpu):
    """128x128 crop, 2 channels, full pyramid (12 levels), 50 iterations;
    loss = masked SSE after registration; derivatives along alpha (both),
    a_data (both channels) and two random moving-frame directions.

    The map is piecewise smooth: every level's 5x5 median and every warp
    sample that crosses the image border put kinks into it, spaced far below
    1e-4 in these directions, so central differences estimate the derivative
    only for h well below that.  Measured convergence (relative error of
    the difference, h = 1e-5 ... 1e-10): alpha converges for h in
    [1e-6, 3e-8], a_data in [1e-7, 3e-9], moving-frame directions in
    [1e-9, 1e-10] (kinks of the map are that close there); every step of
    the sequence taken from the window must agree to 1e-3.  a_data = 0.7: at 0.45 the robust weight
    (E + 1e-5)^(a - 1) is stiff enough (d psi / dE ~ 1e7 where the fit is
    exact) that no step in FP64 isolates a smooth piece; the default exponent
    is covered by the torch autograd comparison and the single-level check."""
    fixed, moving = _frames(2, 128, 128, 2, seed=27)
    mask = np.zeros((128, 128))
    mask[8:-8, 8:-8] = 1.0
    steps = dict(alpha=(3e-7, 1e-7, 3e-8), a_data=(1e-7, 3e-8, 1e-8), moving=(1e-9, 3e-10, 1e-10))
    _assert_fd(_fd_check(fixed, moving, mask, _directions(moving, 13), steps, a_data=0.7))


@pytest.mark.timeout(600)
def test_adjoint_gradient_vs_central_differences_single_level(gpu):
    """levels=0 at the default a_data = 0.45 (one 128x128 level, no warp
    between levels): each of h = 1e-6, 3e-7, 1e-7 for alpha and a_data and
    of h = 1e-9, 3e-10, 1e-10 along moving-frame directions, where the stiff
    robust weight shortens the smooth pieces (measured windows: alpha and
    a_data [3e-6, 1e-8], moving [1e-9, 1e-10])."""
    fixed, moving = _frames(2, 128, 128, 2, seed=28)
    mask = np.zeros((128, 128))
    mask[8:-8, 8:-8] = 1.0
    steps = dict(alpha=(1e-6, 3e-7, 1e-7), a_data=(1e-6, 3e-7, 1e-7), moving=(1e-9, 3e-10, 1e-10))
    _assert_fd(_fd_check(fixed, moving, mask, _directions(moving, 14), steps, levels=0))


@pytest.mark.timeout(900)
def test_gradient_vs_torch_autograd_128(gpu, oracle):
    """The same 128x128 crop, one moving frame, against the torch autograd
    restatement of the whole pyramid."""
    from pyflowreg_amd import autograd as AG
    fixed, moving = _frames(1, 128, 128, 2, seed=27)
    f, m = _planes(fixed, moving)
    mask = torch.zeros((128, 128), dtype=F64)
    mask[8:-8, 8:-8] = 1.0
    alpha_r = torch.tensor([2.0, 2.0], dtype=F64, requires_grad=True)
    a_r = torch.tensor([0.45, 0.45], dtype=F64, requires_grad=True)
    m_r = m.clone().requires_grad_(True)
    flow_r = R.get_displacement(f, m_r, alpha_r, a_r, oracle)
    R.loss(f, m_r, flow_r, mask).backward()
    alpha = torch.tensor([2.0, 2.0], dtype=F64, device="cuda", requires_grad=True)
    a = torch.tensor([0.45, 0.45], dtype=F64, device="cuda", requires_grad=True)
    mv = torch.from_numpy(moving).cuda().requires_grad_(True)
    fx = torch.from_numpy(fixed).cuda()
    flow = AG.get_displacement(fx, mv, alpha=alpha, a_data=a)
    AG.registration_loss(fx, mv, flow, mask.cuda()).backward()
    errs = dict(flow=_rel(flow.permute(0, 3, 1, 2), flow_r), alpha=_rel(alpha.grad, alpha_r.grad),
                a_data=_rel(a.grad, a_r.grad), moving=_rel(mv.grad.permute(0, 3, 1, 2), m_r.grad))
    print("relative errors vs torch autograd:", errs, flush=True)
    # 12 levels x 50 sweeps of the robust solve amplify rounding: ~4e-9
    assert errs["flow"] < 1e-7
    assert errs["alpha"] < 1e-5 and errs["a_data"] < 1e-5 and errs["moving"] < 1e-5
