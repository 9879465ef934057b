"""GPU parity of every operator entry point against the oracle.

The FP64 kernels keep the reference's operation order and are compiled
without FP contraction, so stencil outputs must be bit-identical to the C
restatement (which is itself bit-identical to the reference Fortran, see
test_oracle.py).  Reductions (sums) differ only in summation order.
"""
from __future__ import annotations

import numpy as np
import pytest

from osinco3d_amd import capi
from tests.helpers import rand_fields, solver_code

pytestmark = pytest.mark.gpu

SHAPES = [(9, 11, 13), (12, 10, 70), (16, 12, 8)]   # (nz, ny, nx)
BCS = [(0, 0, 0), (0, 1, 0), (1, 1, 1), (1, 0, 1)]


@pytest.fixture(scope="module", autouse=True)
def gpu():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    capi.load()
    print(f"[gpu] libo3d {capi.load().o3d_version().decode()}", flush=True)
    return torch


def _kind(bc_axis, odd):
    return 0 if bc_axis == 0 else (2 if odd else 1)


@pytest.mark.parametrize("shape", SHAPES)
def test_derivatives_bit_exact(orc, shape):
    f, = rand_fields(shape, 1, seed=3)
    h = 0.173
    for which in range(20):
        df = np.full_like(f, np.nan)
        capi.der(which, df, f, h)
        if which >= 18:
            assert np.array_equal(df, 0.0 * f * h)
            continue
        axis, order, kind = which // 6, (which % 6) // 3 + 1, which % 3
        exp = orc.der1(f, axis, kind, h) if order == 1 else orc.der2(f, axis, kind, h)
        assert np.array_equal(df, exp), (which, np.abs(df - exp).max())


def test_derivatives_device_pointers(orc, gpu):
    torch = gpu
    f, = rand_fields((14, 18, 66), 1, seed=4)
    ft = torch.from_numpy(f).cuda()
    dft = torch.empty_like(ft)
    for which in (0, 5, 7, 11, 14, 15):
        capi.der(which, dft, ft, 0.05)
        torch.cuda.synchronize()
        axis, order, kind = which // 6, (which % 6) // 3 + 1, which % 3
        exp = orc.der1(f, axis, kind, 0.05) if order == 1 else orc.der2(f, axis, kind, 0.05)
        assert np.array_equal(dft.cpu().numpy(), exp)


def test_derivatives_f32_close(orc):
    f, = rand_fields((10, 12, 14), 1, seed=5)
    f32 = f.astype(np.float32)
    for which in range(18):
        df = np.empty_like(f32)
        capi.der(which, df, f32, 0.1)
        axis, order, kind = which // 6, (which % 6) // 3 + 1, which % 3
        exp = orc.der1(f32.astype(np.float64), axis, kind, 0.1) if order == 1 else \
            orc.der2(f32.astype(np.float64), axis, kind, 0.1)
        # FP32 tolerance: a few ulp of the largest term (coefficient * |f|).
        scale = np.abs(exp).max() + (1 / 0.1 if order == 1 else 30 / (12 * 0.01)) * 4
        assert np.abs(df - exp).max() <= 1e-6 * scale


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("odd", [0, 1])
def test_divergence(orc, bc, odd):
    sh = (12, 10, 70)
    fx, fy, fz = rand_fields(sh, 3, seed=7)
    d = (0.11, 0.13, 0.17)
    for divisor in (1.0, 0.37):
        out = np.empty_like(fx)
        st = capi.divergence(out, fx, fy, fz, d, bc, odd, divisor)
        div = orc.divergence(fx, fy, fz, d, bc, odd)
        assert np.array_equal(out, div / divisor)
        ref = orc.function_stats(div)
        assert st[0] == ref[0] and st[1] == ref[1]
        assert np.array_equal(st[3:], ref[3:])
        assert abs(st[2] - ref[2]) <= 1e-13 * np.abs(div).max()


@pytest.mark.parametrize("bc", BCS)
def test_nu_t(orc, bc):
    ux, uy, uz = rand_fields((10, 12, 14), 3, seed=9)
    d = (0.2, 0.21, 0.19)
    out = np.empty_like(ux)
    st = capi.nu_t(out, ux, uy, uz, d, 0.17, 0.2, bc)
    exp = orc.nu_t(ux, uy, uz, d, 0.17, 0.2, bc)
    assert np.array_equal(out, exp)
    ref = orc.function_stats(exp)
    assert st[0] == ref[0] and st[1] == ref[1] and np.array_equal(st[3:], ref[3:])


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("itscheme", [1, 2, 3])
@pytest.mark.parametrize("iles", [0, 1])
def test_predict_velocity(orc, bc, itscheme, iles):
    sh = (10, 12, 14)
    u = rand_fields(sh, 3, seed=11)
    d = (0.11, 0.13, 0.17)
    dt = 0.01
    adt = [dt, 3 * dt / 2, 23 * dt / 12]
    bdt = [0.0, -1 * dt / 2, -16 * dt / 12]
    cdt = [0.0, 0.0, 5 * dt / 12]
    rng = np.random.default_rng(12)
    Fg = [rng.standard_normal((3,) + sh) for _ in range(3)]
    Fo = [f.copy() for f in Fg]
    for itime in (1, 2, 3, 4):
        upg = [np.empty(sh) for _ in range(3)]
        nutg = np.full(sh, np.nan)
        nuto = np.zeros(sh)
        capi.predict_velocity(upg, u, Fg, 1600.0, adt, bdt, cdt, itime, itscheme, d, iles, 0.17,
                              0.2, nutg, bc)
        upo = orc.predict(*u, *Fo, 1600.0, adt, bdt, cdt, itime, itscheme, d, iles, 0.17, 0.2,
                          nuto, bc)
        for a, b in zip(upg, upo):
            assert np.array_equal(a, b), (itime, np.abs(a - b).max())
        for a, b in zip(Fg, Fo):
            assert np.array_equal(a, b)
        assert np.array_equal(nutg, nuto)


def test_predict_rejects_unknown_scheme():
    sh = (8, 8, 8)
    u = rand_fields(sh, 3)
    F = [np.zeros((3,) + sh) for _ in range(3)]
    up = [np.empty(sh) for _ in range(3)]
    with pytest.raises(capi.O3DError, match="itscheme"):
        capi.predict_velocity(up, u, F, 100.0, [1, 1, 1], [0, 0, 0], [0, 0, 0], 5, 7,
                              (0.1, 0.1, 0.1), 0, 0.1, 0.1, np.zeros(sh), (0, 0, 0))


@pytest.mark.parametrize("bc", BCS)
def test_correct_velocity(orc, bc):
    sh = (12, 10, 70)
    up = rand_fields(sh, 3, seed=13)
    pp, = rand_fields(sh, 1, seed=14)
    d = (0.11, 0.13, 0.17)
    ug = [np.empty(sh) for _ in range(3)]
    rc, vs = capi.correct_velocity(ug, up, pp, 0.02, d, bc)
    uo = [np.empty(sh) for _ in range(3)]
    rco = orc.correct_velocity(*uo, *up, pp, 0.02, d, bc)
    assert rc == rco == 0
    for c in range(3):
        assert np.array_equal(ug[c], uo[c])
        assert vs[c, 0] == uo[c].min() and vs[c, 1] == uo[c].max()
        assert vs[c, 2] == np.abs(uo[c]).max() and vs[c, 3] == 0


def test_correct_velocity_guards(orc):
    sh = (8, 8, 8)
    up = rand_fields(sh, 3, seed=15)
    pp = np.zeros(sh)
    d = (0.1, 0.1, 0.1)
    up[1][3, 4, 5] = 1500.0          # maxval(uy) > 1000 -> diverged
    ug = [np.empty(sh) for _ in range(3)]
    rc, vs = capi.correct_velocity(ug, up, pp, 0.02, d, (0, 0, 0))
    assert rc == 1 and orc.correct_velocity(*[np.empty(sh) for _ in range(3)], *up, pp, 0.02, d,
                                            (0, 0, 0)) == 1
    up[1][3, 4, 5] = -1500.0         # signed maximum: a large negative value does not trip it
    rc, _ = capi.correct_velocity(ug, up, pp, 0.02, d, (0, 0, 0))
    assert rc == 0
    up[2][1, 1, 1] = np.nan          # NaN -> diverged
    rc, vs = capi.correct_velocity(ug, up, pp, 0.02, d, (0, 0, 0))
    assert rc == 1 and vs[2, 3] == 1


def test_function_stats(orc):
    f, = rand_fields((13, 17, 71), 1, seed=16)
    st = capi.function_stats(f)
    ref = orc.function_stats(f)
    assert st[0] == ref[0] and st[1] == ref[1] and np.array_equal(st[3:], ref[3:])
    assert abs(st[2] - ref[2]) <= 1e-14 * np.abs(f).max()
    g = np.zeros((9, 9, 9))
    g[2, 3, 4] = 5.0
    g[7, 1, 1] = 5.0                 # tie: the first in column-major order wins
    st = capi.function_stats(g)
    assert list(st[3:]) == [5, 4, 3] == list(orc.function_stats(g)[3:])


@pytest.mark.parametrize("bc", BCS)
def test_statistics(orc, bc):
    u = rand_fields((10, 12, 14), 3, seed=17)
    d = (0.2, 0.21, 0.19)
    g = capi.statistics(*u, d, 1600.0, 1.5, bc)
    o = orc.statistics(*u, d, 1600.0, 1.5, bc)
    assert np.allclose(g, o, rtol=1e-12, atol=0)


# nx = 14, 13, 130: the fused kernel; nx = 16, 20, 12 (multiples of 4): the
# colour-split sweeps.
@pytest.mark.parametrize("solver,shape", [("0000", (10, 12, 14)), ("0011", (10, 9, 14)),
                                          ("111111", (9, 11, 13)), ("0000", (16, 16, 130)),
                                          ("0000", (10, 12, 16)), ("0011", (10, 9, 20)),
                                          ("111111", (9, 11, 12))])
@pytest.mark.parametrize("idyn", [0, 1])
def test_sor_red_black_bit_exact(orc, solver, shape, idyn):
    rhs, p0 = rand_fields(shape, 2, seed=19)
    rhs -= rhs.mean()
    d = (0.3, 0.35, 0.4)
    for eps, kmax in ((1e-8, 400), (1e-30, 7)):
        pg = p0.copy()
        po = p0.copy()
        it, dm, rs, om = capi.poisson(pg, rhs, d, 1.6, eps, kmax, idyn, solver)
        ito, dmo, rso, omo, _ = orc.sor(po, rhs, d, solver, 1.6, eps, kmax, idyn, red_black=True)
        assert (it, rs) == (ito, rso), (it, rs, ito, rso)
        assert dm == dmo and om == omo
        assert np.array_equal(pg, po)


@pytest.mark.parametrize("solver", ["0000", "0011", "111111"])
def test_sor_converges_like_reference(orc, solver):
    """Red-black and lexicographic SOR reach the same field (up to the
    constant null space) at the same eps."""
    sh = (12, 12, 16)
    rhs, = rand_fields(sh, 1, seed=21)
    d = (0.3, 0.3, 0.3)
    # A compatible rhs for every BC: the residual of a random field.
    q, = rand_fields(sh, 1, seed=22)
    rhs = -orc.residual(q, np.zeros(sh), d, solver)
    eps = 1e-11
    pg = np.zeros(sh)
    po = np.zeros(sh)
    capi.poisson(pg, rhs, d, 1.7, eps, 20000, 0, solver)
    orc.sor(po, rhs, d, solver, 1.7, eps, 20000, 0, red_black=False)
    a = pg - pg.mean()
    b = po - po.mean()
    assert np.abs(a - b).max() <= 1e4 * eps
    assert np.abs(a - (q - q.mean())).max() <= 1e4 * eps


def test_sor_rejects_odd_periodic():
    sh = (9, 10, 10)
    with pytest.raises(capi.O3DError, match="even"):
        capi.poisson(np.zeros(sh), np.zeros(sh), (0.1, 0.1, 0.1), 1.5, 1e-8, 10, 0, "0000")


@pytest.mark.parametrize("sh", [(10, 12, 14), (10, 12, 16)])
@pytest.mark.parametrize("bc", [(0, 0, 0), (0, 1, 0), (1, 1, 1)])
def test_correct_pression(orc, bc, sh):
    up = rand_fields(sh, 3, seed=23)
    d = (0.3, 0.32, 0.29)
    dt = 0.05
    p0, = rand_fields(sh, 1, seed=24)
    pg = p0.copy()
    po = p0.copy()
    it, dm, rs, om = capi.correct_pression(pg, up, d, dt, 1.7, 1e-8, 300, 0, bc)
    rhs = orc.divergence(*up, d, bc, 1) / dt
    ito, dmo, rso, omo, _ = orc.sor(po, rhs, d, solver_code(bc), 1.7, 1e-8, 300, 0,
                                    red_black=True)
    assert (it, rs, dm) == (ito, rso, dmo)
    assert np.array_equal(pg, po)


@pytest.mark.parametrize("solver,shape", [("0000", (16, 24, 80)), ("0000", (40, 16, 64)),
                                          ("0011", (12, 17, 72)), ("111111", (9, 11, 16)),
                                          ("0000", (6, 8, 8))])
def test_sor_f32_kernel_variants(solver, shape):
    """The colour-split in-place half sweeps (variant 2) give bit-identical
    iterates (and the same dmax and exit) as the fused one-iteration launch,
    on full and partial tiles, odd y and z counts and every boundary set;
    the removed two-iterations-per-launch variant is refused."""
    nz, ny, nx = shape
    variants = [2]
    with pytest.raises(capi.O3DError, match="variant 1 unknown"):
        capi.poisson_f32(np.zeros(shape, np.float32), np.zeros(shape, np.float32),
                         (0.3, 0.3, 0.3), 1.7, 1e-8, 2, 0, solver, variant=1)
    rhs, p0 = [a.astype(np.float32) for a in rand_fields(shape, 2, seed=41)]
    rhs -= rhs.mean()
    d = (0.3, 0.35, 0.4)
    for kmax in (2, 8, 30):
        ps = p0.copy()
        a = capi.poisson_f32(ps, rhs, d, 1.7, 1e-30, kmax, 0, solver, variant=0)
        for v in variants:
            pp = p0.copy()
            b = capi.poisson_f32(pp, rhs, d, 1.7, 1e-30, kmax, 0, solver, variant=v)
            assert a == b, (v, kmax, a, b)
            assert np.array_equal(ps, pp), (v, kmax, np.abs(ps - pp).max())


def test_sor_f32_converges_to_fp64_solution(orc):
    sh = (16, 16, 24)
    q, = rand_fields(sh, 1, seed=42)
    d = (0.3, 0.3, 0.3)
    rhs = -orc.residual(q, np.zeros(sh), d, "0000")
    p = np.zeros(sh, dtype=np.float32)
    it, dm, rs, _ = capi.poisson_f32(p, rhs.astype(np.float32), d, 1.7, 1e-5, 20000, 0, "0000",
                                     variant=2)
    # eps reached, or the stagnation rule at FP32's rounding floor; never kmax.
    assert rs in (0, 1) and dm < 1e-4
    assert np.abs((p - p.mean()) - (q - q.mean())).max() <= 1e-3


@pytest.mark.parametrize("bc", BCS)
def test_vortical_diagnostics(orc, bc):
    """rotational and calculate_Q_criterion in one fused pass, bit for bit."""
    u = rand_fields((10, 12, 70), 3, seed=43)
    d = (0.11, 0.13, 0.17)
    for a, b in zip(capi.rotational(*u, d, bc), orc.rotational(*u, d, bc)):
        assert np.array_equal(a, b)
    assert np.array_equal(capi.q_criterion(*u, d, bc), orc.q_criterion(*u, d, bc))


@pytest.mark.parametrize("bc", [(0, 0, 0), (0, 1, 0), (1, 1, 1)])
@pytest.mark.parametrize("itscheme,iles", [(1, 0), (2, 1), (3, 0), (3, 1)])
def test_transeq(orc, bc, itscheme, iles):
    """Scalar transport: bit for bit without clipping, round-off with it."""
    from tests.test_oracle import _transeq_case
    sh = (10, 12, 70)
    d = (0.12, 0.14, 0.16)
    dt = 0.01
    adt = [dt, 3 * dt / 2, 23 * dt / 12]
    bdt = [0.0, -dt / 2, -16 * dt / 12]
    cdt = [0.0, 0.0, 5 * dt / 12]
    for clip in (False, True):
        phi, u, src, nut, fphi = _transeq_case(sh, 17 + itscheme, clip)
        p_g, f_g = phi.copy(), fphi.copy()
        p_o, f_o = phi.copy(), fphi.copy()
        for itime in (1, 2, 3):
            capi.transeq(p_g, u, src, f_g, 1000.0, 0.7, adt, bdt, cdt, itime, itscheme, d, iles,
                         nut, bc)
            orc.transeq(p_o, u, src, f_o, 1000.0, 0.7, adt, bdt, cdt, itime, itscheme, d, iles,
                        nut, bc)
            assert np.array_equal(f_g, f_o)
            if clip:
                assert np.abs(p_g - p_o).max() <= 1e-14
                p_g = p_o.copy()
            else:
                assert np.array_equal(p_g, p_o), np.abs(p_g - p_o).max()


def _cc_laplacian(u, h, per):
    """7-point Laplacian of a cell-centred grid (arrays (nz, ny, nx)); axis a
    (x, y, z) periodic when per[a], else the ghost cell equals the edge cell."""
    out = np.zeros_like(u, dtype=np.float64)
    u = u.astype(np.float64)
    for a, ax in ((0, 2), (1, 1), (2, 0)):
        if per[a]:
            s = np.roll(u, 1, ax) + np.roll(u, -1, ax)
        else:
            pad = [(0, 0)] * 3
            pad[ax] = (1, 1)
            up = np.pad(u, pad, mode="edge")
            sl_m = [slice(None)] * 3
            sl_p = [slice(None)] * 3
            sl_m[ax] = slice(0, -2)
            sl_p[ax] = slice(2, None)
            s = up[tuple(sl_m)] + up[tuple(sl_p)]
        out += (s - 2.0 * u) / h[a] ** 2
    return out


@pytest.mark.parametrize("shape,per,h", [((32, 32, 32), (1, 1, 1), (0.1, 0.1, 0.1)),
                                         ((12, 24, 16), (1, 0, 1), (0.1, 0.1, 0.1)),
                                         ((16, 16, 16), (0, 0, 0), (0.1, 0.12, 0.11)),
                                         ((8, 16, 48), (1, 1, 1), (0.05, 0.05, 0.05))])
def test_multigrid_reduces_residual(shape, per, h):
    """The coarse-grid solver of the warm starts: each V(2,2) cycle cuts the
    residual of the cell-centred Poisson problem (checked with an
    independent numpy operator) by at least 5x."""
    f, = rand_fields(shape, 1, seed=51)
    f = (f - f.mean()).astype(np.float32)   # compatible with the null space
    nf = np.linalg.norm(f)
    prev = nf
    for cycles in (1, 2, 3):
        u = np.zeros(shape, dtype=np.float32)
        capi.multigrid(u, f, h, per, cycles)
        r = f - _cc_laplacian(u, h, per)
        r -= r.mean()
        nr = np.linalg.norm(r)
        print(f"[mg] {shape} per {per}: {cycles} cycles, |r|/|f| = {nr / nf:.3e}", flush=True)
        assert nr <= 0.2 ** cycles * nf and nr < prev
        prev = nr


@pytest.mark.parametrize("solver,shape", [("0000", (16, 16, 32)), ("0011", (12, 16, 24)),
                                          ("111111", (16, 12, 16)), ("0000", (8, 10, 40))])
def test_poisson_multigrid_solves_sor_problem(orc, solver, shape):
    """The multigrid solver behind solve_poisson_multigrid's interface
    reaches max |r| < tol on the SOR solvers' own operator (the residual
    recomputed by the oracle), and its field is the exact solution of that
    operator up to the constant null space."""
    q, = rand_fields(shape, 1, seed=61)
    d = (0.3, 0.32, 0.35)
    rhs = -orc.residual(q, np.zeros(shape), d, solver)
    p = np.zeros(shape)
    tol = 1e-9
    cycles, rmax = capi.poisson_multigrid(p, rhs, d, solver, tol=tol, max_cycles=200)
    r = orc.residual(p, rhs, d, solver)
    print(f"[mgp] {solver} {shape}: {cycles} cycles, max|r| {rmax:.3e} (oracle "
          f"{np.abs(r).max():.3e})", flush=True)
    assert rmax < tol and np.abs(r).max() < 1.01 * tol
    assert cycles <= 40
    assert np.abs((p - p.mean()) - (q - q.mean())).max() <= 1e4 * tol


def test_poisson_multigrid_eigenfunction():
    """Periodic eigenfunction of the 7-point operator: L p = -lam p with
    lam = sum (2 - 2 cos(k h)) / h^2 in closed form, so rhs = -lam p has the
    solution p (plus a constant)."""
    nz, ny, nx = 16, 24, 32
    h = (2 * np.pi / nx, 2 * np.pi / ny, 2 * np.pi / nz)
    x = np.arange(nx) * h[0]
    y = np.arange(ny) * h[1]
    z = np.arange(nz) * h[2]
    pe = np.sin(x)[None, None, :] * np.sin(2 * y)[None, :, None] * np.cos(z)[:, None, None]
    lam = sum((2 - 2 * np.cos(k * hh)) / hh ** 2 for k, hh in ((1, h[0]), (2, h[1]), (1, h[2])))
    p = np.zeros_like(pe)
    cycles, rmax = capi.poisson_multigrid(p, -lam * pe, h, "0000", tol=1e-11, max_cycles=100)
    err = np.abs((p - p.mean()) - pe).max()
    print(f"[mgp] eigenfunction: {cycles} cycles, max|r| {rmax:.2e}, error {err:.2e}", flush=True)
    assert rmax < 1e-11 and err < 1e-9


def test_poisson_multigrid_limits():
    """An rhs incompatible with the free-slip mirror: the solve works on its
    compatible part (minus the trapezoid-weighted mean) and ends on tol far
    below a kmax-sized cycle limit, with the oracle's residual of the result
    equal to that mean everywhere; with tol = 0 it runs out of cycles without
    an error, as the SOR solvers only report non-convergence; unsupported
    grids are rejected."""
    sh = (8, 8, 16)
    rng = np.random.default_rng(9)
    rhs = np.ascontiguousarray(1.0 + 0.1 * rng.standard_normal(sh))
    d = (0.1, 0.1, 0.1)
    p = np.zeros(sh)
    cycles, rmax, why = capi.poisson_multigrid(p, rhs, d, "111111", tol=1e-8,
                                               max_cycles=5000, with_reason=True)
    print(f"[mgp] incompatible rhs: {cycles} cycles, max|r| {rmax:.3e}, reason {why}",
          flush=True)
    assert why == 0 and cycles < 60 and rmax < 1e-8
    w = [np.ones(n) for n in sh]
    for a in w:
        a[[0, -1]] = 0.5
    W = w[0][:, None, None] * w[1][None, :, None] * w[2][None, None, :]
    m = (W * rhs).sum() / W.sum()
    from oracle import load_oracle
    r = load_oracle().residual(p, rhs, d, "111111")
    assert np.abs(r - m).max() <= 1e-8
    p = np.zeros(sh)
    cycles, rmax, why = capi.poisson_multigrid(p, rhs, d, "111111", tol=0.0, max_cycles=3,
                                               with_reason=True)
    assert cycles == 3 and why == 2
    with pytest.raises(capi.O3DError, match="multiple of 8"):
        capi.poisson_multigrid(np.zeros((8, 8, 12)), np.zeros((8, 8, 12)), (0.1, 0.1, 0.1),
                               "0000")
    with pytest.raises(capi.O3DError, match="npre"):
        capi.poisson_multigrid(np.zeros(sh), np.zeros(sh), (0.1, 0.1, 0.1), "0000", npre=0,
                               npost=0)


def test_measurement_entry_points():
    """The bench's ceilings: the stream probe (a hand-written triad) reports
    a bandwidth between a copy's plausible floor and the 8 TB/s peak; the
    back-to-back SOR period of a mixed solver is positive, below the
    in-step batches' figure plus slack, and leaves the solver able to step
    on (its buffers are scratch between solves)."""
    from osinco3d_amd.config import tgv_config
    from osinco3d_amd.solver import MIXED, WARM_EXTRAPOLATE, Solver
    gbs, ms = capi.stream_probe(1 << 28, 5)
    assert 1000.0 < gbs < 8000.0 and ms > 0.0
    cfg = tgv_config(64, 6, 3)
    s = Solver(cfg, MIXED, WARM_EXTRAPOLATE)
    s.init_taylor_green()
    for it in range(1, 4):
        s.step(it)
    us = s.sor_period(50)
    assert 0.0 < us < 1000.0
    for it in range(4, 7):
        info = s.step(it)
        assert info["sor_reason"] in (0, 1) and not info["diverged"]
    s.close()
    with pytest.raises(capi.O3DError):
        s2 = Solver(cfg, MIXED)
        try:
            s2.sor_period(1)
        finally:
            s2.close()
