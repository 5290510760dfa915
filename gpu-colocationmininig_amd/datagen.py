"""Seeded synthetic datasets in the reference's CSV format.

Writes what colocationFinder::populateData reads (Codes/colocationFinder.cpp:
139-254): line 1 = N; rows "label,x,y" grouped by integer label 0..F-1 with
float64 coordinates printed so that they round-trip; an empty line; the
bounds line "0,0,S,S".  Also writes parameter.txt ("r theta 1",
cpp:66-68) and config.txt (cpp:1077-1093) next to it.

Workloads (all on a square of side S, coordinates clamped into [0, S)):
  c1  10 features x 2,000 uniform points, S=1e4; 3 planted size-3 patterns,
      60% of each planted feature's instances placed within r/2 of a shared
      seed; r=50, theta=0.3, max size 4.
  c2  500k points, 20 features with Zipf(1.0) frequencies, S=1e5; 70%
      uniform, 30% in 200 Gaussian hotspots (sigma 500); 4 size-3 and 2
      size-4 planted patterns at 50%; r=200, theta=0.25, max size 5.
  c3  5M points, 40 features Zipf(1.0); 80% in 500 hotspots with sigma
      log-uniform in [100, 1000]; 10 planted patterns of size 3-5; r for a
      mean of ~32 neighbours; theta=0.2, max size 5.
  c5  50M points, 60 features Zipf(1.2); 60% in 2,000 hotspots (sigma
      log-uniform in [100, 2000]), each hotspot hosting 12 features; 20
      planted patterns of size 3-6, each a run of consecutive frequency
      ranks in 6..59 (members of similar frequency); r for ~64 neighbours;
      max size 6; thetas 0.1, 0.2, 0.3, 0.5.

Hotspots are shared by all features unless hot_feats is set, in which case
each hotspot hosts hot_feats features drawn at random and a feature's hotspot
points fall only into its own hotspots.  At 50M points with shared hotspots
every feature set co-locates in every hotspot: the prevalent size-3 patterns
alone hold 2.5e10 instances at theta 0.2, and exact mining to size 6 becomes
an enumeration of >1e12 cliques.
"""
import os

import numpy as np

CONFIGS = {
    "c1": dict(seed=1, S=1e4, F=10, N=20_000, zipf=None, hot_frac=0.0, hotspots=0, sigma=(0, 0),
               planted=[3, 3, 3], plant_rate=0.6, r=50.0, theta=0.3, max_k=4),
    "c2": dict(seed=2, S=1e5, F=20, N=500_000, zipf=1.0, hot_frac=0.3, hotspots=200, sigma=(500, 500),
               planted=[3, 3, 3, 3, 4, 4], plant_rate=0.5, r=200.0, theta=0.25, max_k=5),
    "c3": dict(seed=3, S=1e5, F=40, N=5_000_000, zipf=1.0, hot_frac=0.8, hotspots=500, sigma=(100, 1000),
               planted=[3, 3, 3, 4, 4, 4, 5, 5, 3, 4], plant_rate=0.5, r=None, mean_nbrs=32, theta=0.2, max_k=5),
    "c5": dict(seed=5, S=1e5, F=60, N=50_000_000, zipf=1.2, hot_frac=0.6, hotspots=2000, sigma=(100, 2000),
               planted=[3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 6, 6, 6, 6, 3, 4, 5, 6], plant_rate=0.5, r=None,
               mean_nbrs=64, theta=0.1, thetas=[0.1, 0.2, 0.3, 0.5], max_k=6, hot_feats=12, plant_band=(6, 60)),
}


def _feature_counts(N, F, zipf, rng):
    if zipf is None:
        base = np.full(F, N // F, np.int64)
    else:
        w = 1.0 / np.arange(1, F + 1) ** zipf
        base = np.floor(N * w / w.sum()).astype(np.int64)
    base[0] += N - base.sum()
    assert base.min() > 0
    return base


def _points(n, S, hot_frac, centers, sigmas, rng):
    xy = rng.uniform(0, S, size=(n, 2))
    if len(centers) and hot_frac > 0:
        hot = rng.random(n) < hot_frac
        k = int(hot.sum())
        h = rng.integers(0, len(centers), k)
        xy[hot] = centers[h] + rng.normal(size=(k, 2)) * sigmas[h, None]
    return xy


def _clamp_cells(x, y, S, r):
    """Clamps into [0, S) and keeps every point strictly inside the grid as
    the reference computes it in binary32 (Codes/DeviceProcess_Kernel.h:
    615-625), so no point lands on column gx or row gy."""
    top = float(np.nextafter(np.float32(S), np.float32(0)))
    x = np.clip(x, 0.0, top)
    y = np.clip(y, 0.0, top)
    r32 = np.float32(r)
    gx = int(np.ceil(np.float32(S) / r32))
    for a in (x, y):
        i = ((a.astype(np.float32) - np.float32(0)) / r32).astype(np.int64)
        bad = i >= gx
        a[bad] = top - float(r)  # pull back one cell
    return x, y


def training_config(rng, n_range=(100_000, 2_000_000), f_range=(10, 40), wide=False):
    """Randomised parameters of the c4 training datasets (drawn around the
    c1/c2 generator): features in f_range (10-40), instances in n_range,
    hotspot fraction 0-0.6, sigma 0.1-2% of the side, r for 8-64 mean
    neighbours, theta 0.05-0.5, hotspots shared by all features (40%) or
    hosting 2..F features each.  Returns keyword overrides for
    generate("c2", ...).

    wide=True also draws (after the draws above, which it leaves as they
    are) the regimes of the c3/c5 workloads: hotspot fraction 0-0.8, a
    sigma range up to 2% of the side per hotspot, patterns planted up to
    size 6, mined to sizes 4-6 (theta at least 0.1 above size 4), and half
    of the time planted in runs of similar frequency rank (plant_band)."""
    F = int(rng.integers(f_range[0], f_range[1] + 1))
    lo = float(rng.uniform(100, 2000))
    hf = None if rng.random() < 0.4 else int(rng.integers(2, F + 1))
    cfg = dict(F=F, N=int(rng.integers(n_range[0], n_range[1] + 1)), hot_frac=float(rng.uniform(0.0, 0.6)),
               hotspots=int(rng.integers(50, 2001)), sigma=(lo, lo), zipf=float(rng.uniform(0.5, 1.3)),
               planted=[int(v) for v in rng.integers(3, 6, size=int(rng.integers(2, 8)))],
               r=None, mean_nbrs=float(rng.uniform(8, 64)), theta=float(rng.uniform(0.05, 0.5)), max_k=4,
               hot_feats=hf)
    if wide:
        cfg["hot_frac"] = float(rng.uniform(0.0, 0.8))
        cfg["sigma"] = (lo, float(rng.uniform(lo, 2000)))
        cfg["max_k"] = int(rng.integers(4, 7))
        if cfg["max_k"] > 4:  # bounds the candidate count of the deeper levels
            cfg["theta"] = max(cfg["theta"], 0.1)
        cfg["planted"] = [int(v) for v in rng.integers(3, 7, size=int(rng.integers(2, 21)))]
        if rng.random() < 0.5 and F >= 12:
            cfg["plant_band"] = (min(6, F // 4), F)
    return cfg


def generate(name, N=None, seed=None, **over):
    """Returns (labels, x, y, S, params) for workload `name`; N/seed override."""
    cfg = dict(CONFIGS[name])
    cfg.update(over)
    if N is not None:
        cfg["N"] = int(N)
    if seed is not None:
        cfg["seed"] = int(seed)
    rng = np.random.default_rng(cfg["seed"])
    S, F = cfg["S"], cfg["F"]
    counts = _feature_counts(cfg["N"], F, cfg["zipf"], rng)
    nh = cfg["hotspots"]
    centers = rng.uniform(0, S, size=(nh, 2)) if nh else np.zeros((0, 2))
    lo, hi = cfg["sigma"]
    sig = np.exp(rng.uniform(np.log(max(lo, 1e-9)), np.log(max(hi, 1e-9)), nh)) if nh else np.zeros(0)
    if nh and lo == hi:
        sig[:] = lo
    own = [np.arange(nh)] * F
    if cfg.get("hot_feats") and nh:
        hosts = np.array([rng.choice(F, size=cfg["hot_feats"], replace=False) for _ in range(nh)])
        own = [np.nonzero((hosts == f).any(1))[0] for f in range(F)]
        own = [o if len(o) else np.arange(nh) for o in own]
    xs, ys = [], []
    for f in range(F):
        p = _points(int(counts[f]), S, cfg["hot_frac"], centers[own[f]], sig[own[f]], rng)
        xs.append(p[:, 0])
        ys.append(p[:, 1])
    r = cfg["r"]
    if r is None:
        r = estimate_r(np.concatenate(xs), np.concatenate(ys), cfg["mean_nbrs"], rng)
    # planted patterns: for plant_rate of the least frequent member's
    # instances, all members go within r/2 of a shared seed point.  Members
    # are disjoint feature sets drawn at random, or, with plant_band = (lo,
    # hi), runs of consecutive frequency ranks inside [lo, hi): members of
    # similar frequency, so that a planted set's participation index stays
    # near plant_rate for every member (with Zipf frequencies, a set of
    # random members is limited by its most frequent one)
    used = np.zeros(F, np.int64)
    order = rng.permutation(F)
    pos = 0
    planted = []
    band = cfg.get("plant_band")
    taken = np.zeros(F, bool)
    for size in cfg["planted"]:
        if band:
            # a window of free ranks; the band is reused once no window fits
            free = [st for st in range(band[0], band[1] - size + 1) if not taken[st:st + size].any()]
            if not free:
                taken[:] = False
                free = list(range(band[0], band[1] - size + 1))
            start = int(free[int(rng.integers(0, len(free)))])
            taken[start:start + size] = True
            feats = np.arange(start, start + size)
        else:
            if pos + size > F:
                order = rng.permutation(F)
                pos = 0
            feats = np.sort(order[pos:pos + size])
            pos += size
        planted.append(feats.tolist())
        m = int(cfg["plant_rate"] * min(counts[f] - used[f] for f in feats))
        if m <= 0:
            continue
        seeds = _points(m, S, cfg["hot_frac"] * 0.5, centers, sig, rng)
        for f in feats:
            ang = rng.uniform(0, 2 * np.pi, m)
            rad = (r / 2) * np.sqrt(rng.uniform(0, 1, m)) * 0.999
            a, b = used[f], used[f] + m
            xs[f][a:b] = seeds[:, 0] + rad * np.cos(ang)
            ys[f][a:b] = seeds[:, 1] + rad * np.sin(ang)
            used[f] += m
    labels = np.repeat(np.arange(F), counts)
    x, y = _clamp_cells(np.concatenate(xs), np.concatenate(ys), S, r)
    params = dict(r=float(r), theta=cfg["theta"], max_k=cfg["max_k"], S=S, planted=planted,
                  thetas=cfg.get("thetas", [cfg["theta"]]), name=name, N=int(cfg["N"]), F=F, seed=cfg["seed"],
                  zipf=cfg["zipf"], hot_frac=cfg["hot_frac"], hotspots=nh, sigma=tuple(cfg["sigma"]),
                  hot_feats=cfg.get("hot_feats"), plant_rate=cfg["plant_rate"], plant_band=band)
    return labels, x, y, S, params


def estimate_r(x, y, mean_nbrs, rng, sample=20000):
    """r (rounded to 3 significant digits) giving about mean_nbrs neighbours
    per point (including itself) on a random sample.

    Bisection on the radius over the sample's exact mean neighbour count.
    Counts are taken from a k-d tree over only the points near the sample: a
    radius R whose mean count already reaches mean_nbrs is found on a grid of
    R-sized cells, and the tree holds the points of the sample's 3x3 cells.
    Every radius below R is counted exactly; every radius above R has a mean
    count of at least the one at R, so the bisection takes the same branch
    as with a tree over all points."""
    from scipy.spatial import cKDTree
    n = len(x)
    idx = rng.choice(n, size=min(sample, n), replace=False)
    q = np.column_stack([x[idx], y[idx]])
    x0, y0 = float(x.min()), float(y.min())
    area = max((float(x.max()) - x0) * (float(y.max()) - y0), 1e-12)
    R = float(np.sqrt(mean_nbrs * area / (np.pi * n)))  # uniform-density guess
    while True:
        nx = int((float(x.max()) - x0) / R) + 3
        ny = int((float(y.max()) - y0) / R) + 3
        qc = np.floor((q[:, 0] - x0) / R).astype(np.int64) + 1, np.floor((q[:, 1] - y0) / R).astype(np.int64) + 1
        need = np.zeros((ny, nx), bool)
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                need[np.clip(qc[1] + dy, 0, ny - 1), np.clip(qc[0] + dx, 0, nx - 1)] = True
        cx = np.floor((x - x0) / R).astype(np.int64) + 1
        cy = np.floor((y - y0) / R).astype(np.int64) + 1
        sel = need[np.clip(cy, 0, ny - 1), np.clip(cx, 0, nx - 1)]
        del cx, cy
        tree = cKDTree(np.column_stack([x[sel], y[sel]]))
        del sel
        exact = 0.99 * R  # points within this of a sample point lie in its 3x3 cells
        if tree.query_ball_point(q, exact, return_length=True).mean() >= mean_nbrs or R > 1e5:
            break
        R *= 2.0
    lo, hi = 1e-6, 1e5
    for _ in range(40):
        mid = np.sqrt(lo * hi)
        m = tree.query_ball_point(q, mid, return_length=True).mean() if mid <= exact else mean_nbrs
        if m < mean_nbrs:
            lo = mid
        else:
            hi = mid
    return float(f"{np.sqrt(lo * hi):.3g}")


def write_csv(path, labels, x, y, S):
    """The reference dataset format; coordinates as shortest round-trip
    decimals of the float64 values."""
    n = len(x)
    with open(path, "w") as fh:
        fh.write(f"{n}\n")
        CH = 1 << 18
        for s in range(0, n, CH):
            e = min(n, s + CH)
            fh.write("".join(f"{l},{a!r},{b!r}\n" for l, a, b in zip(labels[s:e].tolist(), x[s:e].tolist(), y[s:e].tolist())))
        fh.write("\n")
        fh.write(f"0,0,{S!r},{S!r}")


def write_workload(outdir, name, N=None, seed=None, theta=None, **over):
    """Writes data.csv, parameter.txt, config.txt into outdir; returns paths."""
    os.makedirs(outdir, exist_ok=True)
    labels, x, y, S, p = generate(name, N=N, seed=seed, **over)
    csv = os.path.join(outdir, "data.csv")
    write_csv(csv, labels, x, y, S)
    th = p["theta"] if theta is None else theta
    par = os.path.join(outdir, "parameter.txt")
    with open(par, "w") as fh:
        fh.write(f"{p['r']!r}\n{th!r}\n1\n")
    cfg = os.path.join(outdir, "config.txt")
    with open(cfg, "w") as fh:
        fh.write(f"{outdir}/\n{par}\n{csv}\n{outdir}/\nlog.csv\n")
    p["theta"] = th
    return dict(csv=csv, param=par, config=cfg, params=p)


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("name", choices=sorted(CONFIGS))
    ap.add_argument("outdir")
    ap.add_argument("--N", type=int)
    ap.add_argument("--seed", type=int)
    a = ap.parse_args()
    print(json.dumps(write_workload(a.outdir, a.name, N=a.N, seed=a.seed)["params"]))
