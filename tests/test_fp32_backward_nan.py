"""Cause of the NaN that an earlier FP32 version of the NCC backward passes produced in the summed
pose gradient of the 8-view sphere set, shown on the CPU, and that today's arithmetic is free of it.

The sphere scene's sky is a flat grey of 200.  A source window that lies on the sky has zero
variance, but its FP32 variance E[x^2] - mean^2 cancels two numbers near 4e4 whose spacing is
~0.004, so it comes out as 0 or as one or two of those spacings.  That FP32 backward evaluated it
twice in two forms: its gate used the forward's form of the time, sums divided by sum(w), and found
den = sqrt(var_r var_s) above the degenerate floor 1e-5 * 255^2 (cost = rounding noise near 1, so
the view counted as differentiable); its gradient used sums times 1/sum(w), got var_s = 0, and so
A = 1/den = inf.  On a constant window the bilinear image gradient is 0 and val - mean is 0, and
inf * 0 = NaN entered the per-block pose sums.

Today the forward and pass 1 compute var_s once (one form) and take the reciprocal only above the
floor, so those pairs are degenerate (cost exactly 1) and carry no gradient; the FP64 passes guard
den > 0, so constant windows that the FP32 forward still counts as differentiable (FP32 variance one
spacing above 0) give a zero gradient, not NaN.

The old passes are emulated here in numpy FP32 (fma as an exact product rounded once through FP64,
taps as the forward samples them); the views, windows, homographies and planes come from today's
oracle, whose homography rounds differently from the kernel of that time, so the counts of affected
pairs are illustrative (two reference views of the eight; one selected pair suffices for a NaN sum).
"""
import numpy as np
import pytest

import oracle as O
from lr3d import scenes

f32 = np.float32
FLOOR = f32(1e-5 * 255.0 * 255.0)
REFS = (3, 4)


def fma(a, b, c):
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(f32)


@pytest.fixture(scope="module")
def pairs():
    """Per reference view: window, taps and FP32 sums of every (valid pixel, source view) pair."""
    work = scenes.sphere_small().render()
    imgs = O.Images(work.images)
    W, H = 320, 240
    ox = np.array([(t % 6) * 2 - 5 for t in range(36)], f32)
    oy = np.array([(t // 6) * 2 - 5 for t in range(36)], f32)
    out = []
    for ref in REFS:
        lo, hi = work.ranges[ref]
        srcs = work.sources[ref]
        _, hyp, cost = O.solve(work.views, imgs, ref, srcs, lo, hi, iterations=work.iterations, topk=work.topk)
        p = O.make_params(work.views, ref, srcs, lo, hi, topk=work.topk)
        k = work.views[ref].intrinsics
        idx = np.nonzero((hyp[:, 3] > 0) & (cost < 2.0))[0]
        xs, ys = (idx % W).astype(np.int32), (idx // W).astype(np.int32)
        P, S = idx.size, len(srcs)
        vals = np.empty((P, 36), f32)
        wts = np.empty((P, 36), f32)
        st = np.empty((P, 3), f32)
        Hs = np.empty((P, S, 9), f32)
        for j in range(P):
            x, y = int(xs[j]), int(ys[j])
            vals[j], wts[j], st[j] = O.window(p, imgs, x, y)
            n, d = hyp[idx[j], :3], hyp[idx[j], 3]
            rx = (f32(x) - f32(k.cx)) / f32(k.fx)
            ry = (f32(y) - f32(k.cy)) / f32(k.fy)
            rho = f32(n[0] * (rx * d) + n[1] * (ry * d) + n[2] * d)
            for s in range(S):
                Hs[j, s] = O.homography(p, s, n, rho)
        X = (xs[:, None] + ox[None]).astype(f32)
        Y = (ys[:, None] + oy[None]).astype(f32)
        views = []
        for s in range(S):
            im = work.images[srcs[s]].astype(f32)
            h_, w_ = im.shape
            Hm = Hs[:, s, :]
            q = [fma(Hm[:, 3 * r:3 * r + 1], X, fma(Hm[:, 3 * r + 1:3 * r + 2], Y,
                                                     np.broadcast_to(Hm[:, 3 * r + 2:3 * r + 3], X.shape)))
                 for r in range(3)]
            with np.errstate(all="ignore"):
                rz = (f32(1) / q[2]).astype(f32)
                u = (q[0] * rz).astype(f32)
                v = (q[1] * rz).astype(f32)
            inside = ((q[2] > 1e-6) & (u >= 0) & (v >= 0) & (u <= w_ - 1) & (v <= h_ - 1)).all(1)
            u = np.clip(np.nan_to_num(u), 0, w_ - 1).astype(f32)
            v = np.clip(np.nan_to_num(v), 0, h_ - 1).astype(f32)
            x0, y0 = np.floor(u), np.floor(v)
            fx, fy = (u - x0).astype(f32), (v - y0).astype(f32)
            ix, iy = x0.astype(int), y0.astype(int)
            ix1, iy1 = np.minimum(ix + 1, w_ - 1), np.minimum(iy + 1, h_ - 1)
            a, b, c, e = im[iy, ix], im[iy, ix1], im[iy1, ix], im[iy1, ix1]
            top = fma(fx, (b - a).astype(f32), a)
            bot = fma(fx, (e - c).astype(f32), c)
            val = fma(fy, (bot - top).astype(f32), top)
            gu = ((b - a) * (f32(1) - fy) + (e - c) * fy).astype(f32)
            gv = (bot - top).astype(f32)
            ss = np.zeros(P, f32)
            sss = np.zeros(P, f32)
            srs = np.zeros(P, f32)
            for t in range(36):
                wv = (wts[:, t] * val[:, t]).astype(f32)
                ss = (ss + wv).astype(f32)
                sss = fma(wv, val[:, t], sss)
                srs = fma(wv, vals[:, t], srs)
            views.append(dict(inside=inside, val=val, gu=gu, gv=gv, ss=ss, sss=sss, srs=srs))
        out.append(dict(ref=ref, p=p, imgs=imgs, xs=xs, ys=ys, vals=vals, wts=wts, st=st, Hs=Hs, views=views))
    return out


def old_gate(sv, st):
    """Gate and cost of the old FP32 backward (the forward of the time): sums / sum(w)."""
    sw, mr, var = st[:, 0], st[:, 1], st[:, 2]
    mean = (sv["ss"] / sw).astype(f32)
    var_s = np.maximum((sv["sss"] / sw).astype(f32) - (mean * mean).astype(f32), f32(0))
    den = np.sqrt((var * var_s).astype(f32)).astype(f32)
    cov = ((sv["srs"] / sw).astype(f32) - (mr * mean).astype(f32)).astype(f32)
    with np.errstate(all="ignore"):
        cost = np.where(den < FLOOR, f32(1), np.clip(f32(1) - (cov / den).astype(f32), 0, 2))
    cost = np.where(sv["inside"], cost, f32(2))
    return den, cost


def old_grad_terms(sv, st, vals, wts):
    """dC/dval_i * image gradient of the old FP32 gradient pass: sums * (1 / sum(w)), no guard."""
    sw, mr, var = st[:, 0], st[:, 1], st[:, 2]
    inv_sw = (f32(1) / sw).astype(f32)
    mean = (sv["ss"] * inv_sw).astype(f32)
    var_s = np.maximum((sv["sss"] * inv_sw).astype(f32) - (mean * mean).astype(f32), f32(0))
    den = np.sqrt((var * var_s).astype(f32)).astype(f32)
    cov = ((sv["srs"] * inv_sw).astype(f32) - (mr * mean).astype(f32)).astype(f32)
    with np.errstate(all="ignore"):
        A = (f32(1) / den).astype(f32)
        B = (cov[:, None] * var[:, None] * A[:, None] * A[:, None] * A[:, None]).astype(f32)
        dval = (wts * inv_sw[:, None] * (B * (sv["val"] - mean[:, None]) - (vals - mr[:, None]) * A[:, None])).astype(f32)
        return var_s, dval * sv["gu"], dval * sv["gv"]


def selected(costs, topk):
    """Top-k views by (cost, index) per pixel, as both the forward and pass 1 select them."""
    order = np.argsort(costs, axis=1, kind="stable")[:, :topk]
    sel = np.zeros(costs.shape, bool)
    np.put_along_axis(sel, order, True, axis=1)
    return sel


def test_old_fp32_backward_takes_inf_times_zero_on_sky_windows(pairs):
    bad_total = nan_total = 0
    for r in pairs:
        gates = [old_gate(sv, r["st"]) for sv in r["views"]]
        costs = np.stack([c for _, c in gates], 1)
        sel = selected(costs, 2)
        for s, sv in enumerate(r["views"]):
            den, cost = gates[s]
            smooth = sel[:, s] & (cost > 0) & (cost < 2) & (cost != 1) & ~(den < FLOOR)
            var_s, gx, gy = old_grad_terms(sv, r["st"], r["vals"], r["wts"])
            bad = smooth & (var_s == 0)
            nonfinite = smooth & ~(np.isfinite(gx).all(1) & np.isfinite(gy).all(1))
            # every non-finite gradient of a differentiable pair is one of these gate/gradient
            # mismatches, and every mismatch gives a non-finite gradient
            assert np.array_equal(bad, nonfinite)
            # they are sky windows: every tap reads the flat grey
            assert (sv["val"][bad] == 200.0).all()
            bad_total += int(bad.sum())
            nan_total += int(np.isnan(gx[bad]).any(1).sum() + np.isnan(gy[bad]).any(1).sum() > 0)
    print(f"selected differentiable pairs with var_s > 0 at the gate and 0 in the gradient: {bad_total}")
    assert bad_total > 0 and nan_total > 0


def test_todays_fp32_cost_has_one_form_and_skips_those_pairs(pairs):
    """The oracle's view cost (the forward's arithmetic, shared by pass 1 of the adjoint) returns a
    finite cost in [0, 2] on every pair; at the old gate/gradient mismatches it returns exactly 1
    (degenerate), which the adjoint's gate excludes."""
    checked = 0
    for r in pairs:
        for s, sv in enumerate(r["views"]):
            den, cost = old_gate(sv, r["st"])
            var_s, _, _ = old_grad_terms(sv, r["st"], r["vals"], r["wts"])
            bad = np.nonzero(sv["inside"] & ~(den < FLOOR) & (cost > 0) & (cost < 2) & (var_s == 0))[0]
            rng = np.random.default_rng(r["ref"] * 8 + s)
            rest = rng.choice(np.nonzero(sv["inside"])[0], size=200, replace=False)
            for j in np.concatenate([bad, rest]):
                win = (r["vals"][j], r["wts"][j], r["st"][j])
                c = O.view_cost(r["p"], r["imgs"], s, r["Hs"][j, s], int(r["xs"][j]), int(r["ys"][j]), win)
                assert np.isfinite(c) and 0.0 <= c <= 2.0
                if j in set(bad.tolist()):
                    assert c == 1.0
                checked += 1
    assert checked > 0


def test_fp64_passes_give_zero_gradient_on_constant_windows(pairs):
    """Constant source windows that today's FP32 forward still counts as differentiable (FP32
    variance one spacing above zero): the FP64 gradient formula of the backward passes, with its
    den > 0 guard, is finite there and zero (no image gradient on a flat window)."""
    seen = 0
    for r in pairs:
        sw = r["st"][:, 0].astype(np.float64)
        mr = r["st"][:, 1].astype(np.float64)
        var = r["st"][:, 2].astype(np.float64)
        for sv in r["views"]:
            inv_sw = 1.0 / sw
            isw32 = (f32(1) / r["st"][:, 0]).astype(f32)
            mean32 = (sv["ss"] * isw32).astype(f32)
            var32 = np.maximum((sv["sss"] * isw32).astype(f32) - (mean32 * mean32).astype(f32), f32(0))
            den32 = np.sqrt((r["st"][:, 2] * var32).astype(f32))
            const = (sv["val"] == sv["val"][:, :1]).all(1) & (sv["gu"] == 0).all(1) & (sv["gv"] == 0).all(1)
            live = sv["inside"] & const & (den32 >= FLOOR)
            if not live.any():
                continue
            val = sv["val"][live].astype(np.float64)
            w = r["wts"][live].astype(np.float64)
            v = r["vals"][live].astype(np.float64)
            ss = (w * val).sum(1)
            sss = (w * val * val).sum(1)
            srs = (w * val * v).sum(1)
            mean = ss * inv_sw[live]
            var_s = np.maximum(sss * inv_sw[live] - mean * mean, 0.0)
            den = np.sqrt(var[live] * var_s)
            cov = srs * inv_sw[live] - mr[live] * mean
            with np.errstate(divide="ignore"):
                A = np.where(den > 0.0, 1.0 / np.where(den > 0.0, den, 1.0), 0.0)
            B = cov * var[live] * A ** 3
            dval = w * inv_sw[live][:, None] * (B[:, None] * (val - mean[:, None]) - (v - mr[live][:, None]) * A[:, None])
            gx = dval * sv["gu"][live]
            gy = dval * sv["gv"][live]
            assert np.isfinite(gx).all() and np.isfinite(gy).all()
            assert (gx == 0).all() and (gy == 0).all()
            seen += int(live.sum())
    print(f"constant windows the FP32 forward counts as non-degenerate: {seen}")
    assert seen > 0
