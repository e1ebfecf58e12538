"""The FP16 cull of the mixed-precision density pass, emulated on the CPU
operation by operation (every FP16 add, multiply and fused multiply-add is
computed exactly in double and rounded once to FP16, as the packed FP16
instructions round), against the FP32 neighbour test it must never
contradict.

Pairs are drawn so that many lie within a few per cent of r = h, across
cubic, anisotropic and larger-than-h cells, simulation scales other than 1,
particles outside the grid (projected onto their clamped cells) and all 27
cell offsets, including the x offsets whose cell indices differ by 2 mod 3.
Every pair the FP32 test accepts must survive the cull.

Two planted mistakes in the emulation must each drop accepted pairs, so
the draw reaches the cases the margin and the periodic x distance exist for:
no margin on the cut, and no min(a, P - a) in the x distance."""
import numpy as np

F32 = np.float32


def rn16(x):
    return np.asarray(x, np.float64).astype(np.float16)


def f16(x):
    return np.asarray(x, np.float16).astype(np.float64)


def cell_of(p, vmin, cs, dims):
    q = np.floor((p.astype(F32) - F32(vmin)).astype(F32) / F32(cs)).astype(F32)
    return np.clip(q, 0, dims - 1).astype(np.int64)


def offsets(p, vmin, cs, dims, L):
    """(X, Y, Z) in FP32 for positions p (n, 3)."""
    out = np.empty_like(p, dtype=F32)
    c = np.empty(p.shape, np.int64)
    for a in range(3):
        c[:, a] = cell_of(p[:, a], vmin[a], cs[a], dims[a])
        lo = (F32(vmin[a]) + c[:, a].astype(F32) * F32(cs[a])).astype(F32)
        inv = F32(1.0) / F32(cs[a])
        u = np.clip(((p[:, a].astype(F32) - lo).astype(F32) * inv).astype(F32), F32(0), F32(1))
        w = F32(F32(cs[a]) / F32(L))
        if a == 0:
            v = ((u + (c[:, a] % 3).astype(F32)).astype(F32) - F32(1.5)).astype(F32)
        else:
            v = (u - F32(0.5)).astype(F32)
        out[:, a] = (w * v).astype(F32)
    return out, c


def fp32_accepts(pi, pj, hhu):
    d = (pi.astype(F32) - pj.astype(F32)).astype(F32)
    r2 = (d[:, 0] * d[:, 0]).astype(F32)
    r2 = (d[:, 1].astype(np.float64) * d[:, 1] + r2).astype(F32)  # fmaf: one rounding
    r2 = (d[:, 2].astype(np.float64) * d[:, 2] + r2).astype(F32)
    return r2 < F32(hhu)


def cull_survives(oi, oj, dyz, cs, L, hc, margin=2.0 ** -6, wrap=True):
    """The kernel's cull_pair for candidate j of row (dy, dz) of particle i."""
    w = (np.asarray(cs, F32) / F32(L)).astype(F32)
    cut = F32((F32(hc) * F32(hc) + F32(margin)) * F32(1.001953125))
    tx = f16(oi[:, 0])
    ty = f16((oi[:, 1] - (dyz[:, 0].astype(F32) * w[1]).astype(F32)).astype(F32))
    tz = f16((oi[:, 2] - (dyz[:, 1].astype(F32) * w[2]).astype(F32)).astype(F32))
    X, Y, Z = f16(oj[:, 0]), f16(oj[:, 1]), f16(oj[:, 2])
    per = f16(F32(3.0) * w[0])
    a = np.abs(f16(X - tx))
    dx = np.minimum(a, f16(per - a)) if wrap else a
    dy, dz = f16(Y - ty), f16(Z - tz)
    ncut = f16(-cut)
    r = f16(dx * dx + ncut)
    r = f16(dy * dy + r)
    r = rn16(dz * dz + r)
    return np.signbit(r)


def draw(rng, cs, dims, vmin, h_over_s, n):
    """Particles i anywhere in (and around) the grid; candidates j at a
    distance concentrated near h in a random direction.  Keeps the pairs
    whose cells are within one cell on every axis (the stencil)."""
    cs = np.asarray(cs, np.float64)
    ext = cs * dims
    pi = vmin + rng.uniform(-0.08, 1.08, (n, 3)) * ext
    # half the pairs on cell faces (x_i at a face, j across it)
    k = rng.random(n) < 0.5
    ax = rng.integers(0, 3, n)
    face = vmin[ax] + cs[ax] * rng.integers(0, dims[ax] + 1, n)
    pi[k, ax[k]] = face[k] + rng.uniform(-1e-3, 1e-3, k.sum()) * cs[ax[k]]
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    sel = rng.random(n)
    r = h_over_s * np.where(sel < 0.6, 1.0 + rng.normal(0, 2e-3, n),
                            np.where(sel < 0.85, rng.uniform(0, 1.05, n), rng.uniform(1.05, 2.0, n)))
    pj = pi + v * r[:, None]
    return pi.astype(F32), pj.astype(F32)


CASES = [
    # cell size per axis (position units), h/s as a fraction of the smallest
    # cell edge (h/s is the neighbour radius in position units: the second
    # last case is the reference application's, h = 0.018 at scale 0.009)
    ((0.05, 0.05, 0.05), 1.0),
    ((0.025, 0.025, 0.025), 1.0),
    ((0.05, 0.065, 0.085), 1.0),
    ((0.08, 0.08, 0.08), 0.625),
    ((0.07, 0.05, 0.06), 0.99),
    ((2.0, 2.0, 2.0), 1.0),
    ((0.003, 0.003, 0.003), 1.0),
]


def _run(margin=2.0 ** -6, wrap=True, n=200_000, seed=5):
    dropped, accepted, survivors_far, far = 0, 0, 0, 0
    rng = np.random.default_rng(seed)
    for cs, frac in CASES:
        cs32 = tuple(float(F32(c)) for c in cs)
        dims = np.array([40, 30, 35])
        vmin = np.array([-0.3, 0.1, 0.0]) * np.array(cs32) * 10
        h_over_s = float(F32(frac * min(cs32)))
        L = max(cs32)
        pi, pj = draw(rng, cs32, dims, vmin, h_over_s, n)
        oi, ci = offsets(pi, vmin, cs32, dims, L)
        oj, cj = offsets(pj, vmin, cs32, dims, L)
        inrow = np.all(np.abs(cj - ci) <= 1, axis=1)
        oi, oj, ci, cj, pi_, pj_ = oi[inrow], oj[inrow], ci[inrow], cj[inrow], pi[inrow], pj[inrow]
        dyz = (cj - ci)[:, 1:]
        hhu = F32(F32(h_over_s) * F32(h_over_s))
        acc = fp32_accepts(pi_, pj_, hhu)
        surv = cull_survives(oi, oj, dyz, cs32, L, h_over_s / L, margin=margin, wrap=wrap)
        dropped += int(np.count_nonzero(acc & ~surv))
        accepted += int(np.count_nonzero(acc))
        d = np.linalg.norm(pi_.astype(np.float64) - pj_, axis=1)
        hi = vmin + np.asarray(cs32) * dims
        inside = np.all((pi_ >= vmin) & (pi_ <= hi) & (pj_ >= vmin) & (pj_ <= hi), axis=1)
        farm = (d > 1.05 * h_over_s) & inside
        far += int(np.count_nonzero(farm))
        survivors_far += int(np.count_nonzero(surv & farm))
    return dropped, accepted, survivors_far, far


def test_cull_never_drops_an_accepted_pair():
    dropped, accepted, sf, far = _run()
    print(f"accepted {accepted:,}, dropped {dropped}, survivors beyond 1.05 h: {sf:,} of {far:,}")
    assert accepted > 500_000
    assert dropped == 0
    # the cull is still a cull: of the pairs inside the grid (projection
    # leaves them where they are) beyond 1.05 h, none survives
    assert far > 50_000 and sf == 0


def test_planted_mistakes_are_caught():
    d0, _, _, _ = _run(margin=0.0, n=100_000, seed=6)
    d1, _, _, _ = _run(wrap=False, n=100_000, seed=7)
    print(f"no margin: {d0} dropped; no periodic x distance: {d1} dropped")
    assert d0 > 0 and d1 > 0
