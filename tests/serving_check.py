"""Per-job check of a serving-engine image against the CPU FP64 oracle (test tooling; the oracle
is only the checker): `sample` uniform random pixels plus the (2 half + 1)^2 window around the
brightest pixel of the job's image, formed by the oracle from the job's own inputs (range profiles
and accumulation in FP64, oracle/sarbp_oracle.c) under the job's Taylor-ratio policy."""
import numpy as np


def check_job(sar, orc, fac, i, img, sample=128, half=4, seed=7000):
    """Returns dict(rel_l2, rel_l2_peak_window, peak_moved) for job i of JobFactory `fac`."""
    sc, g, sp = fac.scenes[i], fac.grids[i], fac.specs[i]
    side = sp["side"]
    rng = np.random.default_rng(seed + i)
    ys = list(rng.integers(0, side, sample))
    xs = list(rng.integers(0, side, sample))
    py, px = np.unravel_index(int(np.argmax(np.abs(img))), img.shape)
    win = [(y, x) for y in range(py - half, py + half + 1) for x in range(px - half, px + half + 1)
           if 0 <= y < side and 0 <= x < side]
    ys += [w[0] for w in win]
    xs += [w[1] for w in win]
    ys, xs = np.array(ys, np.int32), np.array(xs, np.int32)
    grid = sar.image_grid(side, side, g["d_u"], g["d_u"], g["u_hat"], g["v_hat"],
                          with_pixel_locs=False)
    pts = sar.grid_points(grid, xs, ys)
    cnt = len(xs)
    den = fac.taylor_den_policy
    if den == sar.POLICY_DEFAULT:
        den = sar.get_default_policies()[1]
    Q = orc.range_profiles(fac.phs[i], sc.k_r, sp["taylor"], sar.nfft(sc.nsamples, 2),
                           int_den=den == sar.TAYLOR_DEN_INT)
    _, nr0 = orc.norm_r0(sc.pos)   # the engine's jobs use the default strict f32 ||pos||
    ref = orc.backproject_points(Q, sc.nsamples, sc.delta_r, sc.k_r[sc.nsamples // 2], sc.pos,
                                 nr0, pts, cnt, 1, np.arange(cnt, dtype=np.int32),
                                 np.zeros(cnt, np.int32))
    got = img[ys, xs]
    m = sample
    rel = float(np.linalg.norm(got[:m] - ref[:m]) / np.linalg.norm(ref[:m]))
    relw = float(np.linalg.norm(got[m:] - ref[m:]) / np.linalg.norm(ref[m:]))
    moved = int(np.argmax(np.abs(got[m:])) != np.argmax(np.abs(ref[m:])))
    return dict(rel_l2=rel, rel_l2_peak_window=relw, peak_moved=moved)
