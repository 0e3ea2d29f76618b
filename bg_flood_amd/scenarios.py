"""Synthetic inputs in the reference's BUQ layout for the benchmark workloads.

A scenario is a dict with the Param keys (see bg_flood_amd.default_param),
the block list (BlockP arrays) and the initial arrays zb, h, u, v, cf, each
nblkmem*324 long in memloc order (src/MemManagement.cu:37-40).  Initial
water follows coldstart (src/InitEvolv.cu:131-165): zs = max(zsinit, zb),
h = max(zs - zb, 0), u = v = 0.
"""
import math

import numpy as np

from . import BS, BW, MW, default_param, mesh_uniform


def nblkmem_for(nblk, membuffer=1.05):
    """nblkmem = ceil(nblk * membuffer) (src/Mesh.cu:145)."""
    return int(math.ceil(nblk * membuffer))


def buq_from_grid(field, nblkmem, dtype):
    """Scatters an (ny, nx) raster into BUQ interior cells of a raster-ordered uniform block list."""
    ny, nx = field.shape
    nby, nbx = ny // BW, nx // BW
    out = np.zeros((nblkmem, MW, MW), dtype)
    blocks = field.reshape(nby, BW, nbx, BW).transpose(0, 2, 1, 3).reshape(nby * nbx, BW, BW)
    out[:nby * nbx, 1:MW - 1, 1:MW - 1] = blocks
    return out.reshape(-1)


def grid_from_buq(arr, nbx, nby):
    """Inverse of buq_from_grid for the interior cells."""
    a = np.asarray(arr).reshape(-1, MW, MW)[:nbx * nby, 1:MW - 1, 1:MW - 1]
    return a.reshape(nby, nbx, BW, BW).transpose(0, 2, 1, 3).reshape(nby * BW, nbx * BW)


def cell_coords(blocks, nblkmem, xo, yo, delta, dtype):
    """x = xo + blk.xo + ix*delta, y likewise, evaluated in T like the reference tests."""
    T = dtype
    nb = len(blocks["active"])
    ix = np.arange(BW, dtype=T)
    X = np.zeros((nblkmem, MW, MW), T)
    Y = np.zeros((nblkmem, MW, MW), T)
    bxo = np.asarray(blocks["xo"][:nb]).astype(T)
    byo = np.asarray(blocks["yo"][:nb]).astype(T)
    X[:nb, 1:MW - 1, 1:MW - 1] = (T(xo) + bxo[:, None, None]) + ix[None, None, :] * T(delta)
    Y[:nb, 1:MW - 1, 1:MW - 1] = (T(yo) + byo[:, None, None]) + ix[None, :, None] * T(delta)
    return X.reshape(-1), Y.reshape(-1)


def cell_coords_levels(blocks, nblkmem, xo, yo, dx, dtype):
    """Cell centres of a multi-level block list: x = xo + blk.xo + ix*calcres(dx, level), in T."""
    T = dtype
    nb = min(len(blocks["level"]), nblkmem)  # every slot (active or not) has a level and an origin
    lev = np.asarray(blocks["level"][:nb])
    d = np.where(lev < 0, dx * np.exp2(-lev.astype(np.float64)), dx / np.exp2(lev.astype(np.float64))).astype(T)
    ix = np.arange(BW, dtype=T)
    X = np.zeros((nblkmem, MW, MW), T)
    Y = np.zeros((nblkmem, MW, MW), T)
    bxo = np.asarray(blocks["xo"][:nb]).astype(T)
    byo = np.asarray(blocks["yo"][:nb]).astype(T)
    X[:nb, 1:MW - 1, 1:MW - 1] = (T(xo) + bxo[:, None, None]) + ix[None, None, :] * d[:, None, None]
    Y[:nb, 1:MW - 1, 1:MW - 1] = (T(yo) + byo[:, None, None]) + ix[None, :, None] * d[:, None, None]
    return X.reshape(-1), Y.reshape(-1)


def quadtree_leaves(nbx, nby, maxlevel, refine):
    """Leaves of a 2:1-balanced quadtree over nbx x nby level-0 blocks.

    refine(level, ix, iy) says whether block (ix, iy) of that level is split; blocks are split
    level by level up to maxlevel, and coarse leaves next to leaves two levels finer are split
    until neighbouring levels differ by one at most.  Returns (level, ix, iy) tuples in Z order
    (level-0 blocks in raster order, children bottom-left, bottom-right, top-left, top-right)."""
    leaves = {(0, x, y) for y in range(nby) for x in range(nbx)}

    def split(b):
        l, x, y = b
        leaves.discard(b)
        leaves.update({(l + 1, 2 * x + i, 2 * y + j) for j in (0, 1) for i in (0, 1)})

    def cover(l, x, y):
        for lv in range(l, -1, -1):
            b = (lv, x >> (l - lv), y >> (l - lv))
            if b in leaves:
                return b
        return None

    def balance():
        changed = True
        while changed:
            changed = False
            for (l, x, y) in sorted(leaves, reverse=True):
                if l < 2 or (l, x, y) not in leaves:
                    continue
                for ox, oy in ((-1, 0), (1, 0), (0, -1), (0, 1)):
                    nx_, ny_ = x + ox, y + oy
                    if nx_ < 0 or ny_ < 0 or nx_ >= (nbx << l) or ny_ >= (nby << l):
                        continue
                    c = cover(l, nx_, ny_)
                    if c is not None and c[0] < l - 1:
                        split(c)
                        changed = True

    for l in range(maxlevel):
        for b in sorted(leaves):
            if b[0] == l and refine(*b):
                split(b)
        balance()

    def zkey(b):
        l, x, y = b
        sh = maxlevel - l
        X, Y = x << sh, y << sh
        root = (Y >> maxlevel) * nbx + (X >> maxlevel)
        code = 0
        for k in range(maxlevel):
            code |= ((X >> k) & 1) << (2 * k) | ((Y >> k) & 1) << (2 * k + 1)
        return root, code

    return sorted(leaves, key=zkey)


def uniform(nx, ny, dx, precision="fp64", xo=0.0, yo=0.0, **kw):
    """Param + block list of a uniform one-level nx x ny grid (multiples of 16)."""
    assert nx % BW == 0 and ny % BW == 0
    nbx, nby = nx // BW, ny // BW
    dbl = precision == "fp64"
    blocks = mesh_uniform(nbx, nby, dx, 0, dbl)
    nblk = nbx * nby
    p = default_param(dx=dx, delta=dx, xo=xo, yo=yo, nblk=nblk, nblkmem=nblkmem_for(nblk), precision=precision,
                      **kw)
    return p, blocks, nbx, nby


def gaussian_hump(n=256, dx=10.0, depth=10.0, amp=1.0, radius=200.0, precision="fp64", aoibnd=0):
    """Uniform n x n grid, flat bed zb=-depth, zs = amp*exp(-r^2/(2*radius^2)) at the centre,
    u = v = 0, frictionless, walls: the small plain-method case."""
    p, blocks, nbx, nby = uniform(n, n, dx, precision, aoibnd=aoibnd, frictionmodel=0, cf=0.0)
    T = np.float64 if precision == "fp64" else np.float32
    S = p["nblkmem"] * BS
    X, Y = cell_coords(blocks, p["nblkmem"], 0.0, 0.0, dx, T)
    xc = T(n * dx / 2.0)
    zb = np.full(S, T(-depth), T)
    zs = np.maximum(T(0.0), zb) + T(amp) * np.exp(T(-1.0) * ((X - xc) * (X - xc) + (Y - xc) * (Y - xc))
                                                   / (T(2.0) * T(radius) * T(radius)))
    h = np.maximum(zs - zb, T(0.0)).astype(T)
    inter = np.zeros((p["nblkmem"], MW, MW), bool)
    inter[:p["nblk"], 1:MW - 1, 1:MW - 1] = True
    h[~inter.reshape(-1)] = 0
    z = np.zeros(S, T)
    return dict(param=p, blocks=blocks, nbx=nbx, nby=nby, zb=zb, h=h, u=z.copy(), v=z.copy(), cf=z.copy())


def reference_hump(precision="fp64", aoibnd=1):
    """The reference's Gaussian hump verification setup (src/Testing.cu:418-600): 256^2 cells of
    dx = 1/256 on [-0.5,0.5]^2, zb = 0, zsinit = 0.1, bump a = 1, cc = 0.05, frictionless,
    run to t = 0.1 s."""
    dx = 1.0 / 256
    p, blocks, nbx, nby = uniform(256, 256, dx, precision, xo=-0.5, yo=-0.5, aoibnd=aoibnd, frictionmodel=0,
                                  cf=0.0)
    T = np.float64 if precision == "fp64" else np.float32
    S = p["nblkmem"] * BS
    X, Y = cell_coords(blocks, p["nblkmem"], -0.5, -0.5, dx, T)
    zb = np.zeros(S, T)
    cc = T(0.05)
    zs = np.maximum(T(0.1) + T(0.0), zb) + T(1.0) * np.exp(T(-1.0) * (X * X + Y * Y) / (T(2.0) * cc * cc))
    h = np.maximum(zs - zb, T(0.0)).astype(T)
    z = np.zeros(S, T)
    return dict(param=p, blocks=blocks, nbx=nbx, nby=nby, zb=zb, h=h, u=z.copy(), v=z.copy(), cf=z.copy(),
                nextoutputtime=0.1)


def value_noise(nx, ny, seed, octaves=6, base=4, lacunarity=2, gain=0.5):
    """Sum of `octaves` bilinearly interpolated seeded lattices (base x base cells across the
    domain at the first octave), normalised to [0, 1]."""
    rng = np.random.default_rng(seed)
    out = np.zeros((ny, nx), np.float64)
    amp = 1.0
    cells = base
    ys = (np.arange(ny) + 0.5) / ny
    xs = (np.arange(nx) + 0.5) / nx
    for _ in range(octaves):
        lat = rng.random((cells + 1, cells + 1))
        fy = ys * cells
        fx = xs * cells
        iy = np.minimum(fy.astype(np.int64), cells - 1)
        ix = np.minimum(fx.astype(np.int64), cells - 1)
        ty = (fy - iy)[:, None]
        tx = (fx - ix)[None, :]
        # bilinear interpolation, separable: rows first (ny x cells+1), then columns
        rows = lat[iy] * (1 - ty) + lat[iy + 1] * ty
        if nx % cells == 0:
            rep = nx // cells
            left = np.repeat(rows[:, :-1], rep, axis=1)
            right = np.repeat(rows[:, 1:], rep, axis=1)
        else:
            left, right = rows[:, ix], rows[:, ix + 1]
        left *= amp * (1 - tx)
        right *= amp * tx
        out += left
        out += right
        del left, right
        amp *= gain
        cells *= lacunarity
    out -= out.min()
    out /= max(out.max(), 1e-300)
    return out


def rain_terrain(n=8192, dx=5.0, amp=20.0, slope=0.002, rain=50.0, manning=0.03, seed=1, precision="fp32",
                 aoibnd=1):
    """Rain on a seeded value-noise bed plus a planar slope, initially dry, Manning friction,
    outflow (Neumann) boundaries: the HBM-scale wet/dry workload."""
    p, blocks, nbx, nby = uniform(n, n, dx, precision, aoibnd=aoibnd, frictionmodel=-1, cf=manning, rain_on=1,
                                  rain=rain)
    T = np.float64 if precision == "fp64" else np.float32
    bed = amp * value_noise(n, n, seed)
    x = (np.arange(n) + 0.5) * dx
    bed = bed + slope * x[None, :]
    zb = buq_from_grid(bed.astype(T), p["nblkmem"], T)
    del bed
    S = p["nblkmem"] * BS
    h = np.zeros(S, T)
    cf = np.full(S, T(manning), T)
    return dict(param=p, blocks=blocks, nbx=nbx, nby=nby, zb=zb, h=h, u=np.zeros(S, T), v=np.zeros(S, T), cf=cf)


def lake_at_rest(n=64, dx=1.0, level=0.5, amp=1.0, seed=3, precision="fp64", aoibnd=0):
    """Still water at zs = level over a rough bed that emerges above it (wet/dry)."""
    p, blocks, nbx, nby = uniform(n, n, dx, precision, aoibnd=aoibnd, frictionmodel=0, cf=0.0)
    T = np.float64 if precision == "fp64" else np.float32
    bed = amp * value_noise(n, n, seed, octaves=4) - 0.2
    zb = buq_from_grid(bed.astype(T), p["nblkmem"], T)
    S = p["nblkmem"] * BS
    h = np.maximum(T(level) - zb, T(0)).astype(T)
    inter = np.zeros((p["nblkmem"], MW, MW), bool)
    inter[:p["nblk"], 1:MW - 1, 1:MW - 1] = True
    h[~inter.reshape(-1)] = 0
    z = np.zeros(S, T)
    return dict(param=p, blocks=blocks, nbx=nbx, nby=nby, zb=zb, h=h, u=z.copy(), v=z.copy(), cf=z.copy())


def amr_basin(nbx=4, nby=4, dx=10.0, maxlevel=2, depth=10.0, amp=1.0, radius=None, precision="fp64", aoibnd=0,
              island=False, rain=0.0, frictionmodel=0, cf=0.0, zsinit=0.0, refine=None):
    """A quadtree mesh over nbx x nby level-0 blocks refined towards the centre (maxlevel levels).

    Flat bed zb = -depth; optionally a conical island (peak 2 m above the still level) straddling
    the refinement boundary so wet/dry fronts cross coarse/fine interfaces.  zs = zsinit + a
    Gaussian hump of amplitude amp at the centre where the bed is below zsinit (cold start:
    h = max(zs - zb, 0), u = v = 0)."""
    from . import mesh_from_leaves
    Lx, Ly = nbx * BW * dx, nby * BW * dx
    cx, cy = 0.5 * Lx, 0.5 * Ly
    if refine is None:
        def refine(l, x, y):
            s = BW * dx / (1 << l)
            bx, by = (x + 0.5) * s, (y + 0.5) * s
            return math.hypot(bx - cx, by - cy) < (0.3 if l == 0 else 0.18) * min(Lx, Ly)
    leaves = quadtree_leaves(nbx, nby, maxlevel, refine)
    blocks = mesh_from_leaves(leaves, nbx, nby, dx)
    nblk = len(leaves)
    levels = [b[0] for b in leaves]
    p = default_param(dx=dx, delta=dx, nblk=nblk, nblkmem=nblkmem_for(nblk), precision=precision,
                      minlevel=min(levels), maxlevel=max(levels), aoibnd=aoibnd, frictionmodel=frictionmodel,
                      cf=cf, rain_on=1 if rain else 0, rain=rain)
    T = np.float64 if precision == "fp64" else np.float32
    S = p["nblkmem"] * BS
    X, Y = cell_coords_levels(blocks, p["nblkmem"], 0.0, 0.0, dx, T)
    zb = np.full(S, T(-depth), T)
    if island:
        ix_, iy_ = T(0.32 * Lx), T(0.5 * Ly)
        r = np.sqrt((X - ix_) * (X - ix_) + (Y - iy_) * (Y - iy_))
        zb = np.maximum(zb, (T(2.0) - r * T((depth + 2.0) / (0.2 * Lx))).astype(T)).astype(T)
    rad = T(radius if radius is not None else 0.12 * Lx)
    hump = T(amp) * np.exp(T(-1.0) * ((X - T(cx)) * (X - T(cx)) + (Y - T(cy)) * (Y - T(cy))) / (T(2.0) * rad * rad))
    zs = np.maximum(T(zsinit), zb) + np.where(zb < T(zsinit), hump, T(0.0))
    h = np.maximum(zs - zb, T(0.0)).astype(T)
    inter = np.zeros((p["nblkmem"], MW, MW), bool)
    inter[:nblk, 1:MW - 1, 1:MW - 1] = True
    h[~inter.reshape(-1)] = 0
    z = np.zeros(S, T)
    cfa = np.full(S, T(cf), T)
    return dict(param=p, blocks=blocks, nbx=nbx, nby=nby, zb=zb, h=h, u=z.copy(), v=z.copy(), cf=cfa, leaves=leaves)


def ensemble_specs(n=64, seed=7, lmin=256, lmax=2048):
    """Seeded member specifications of the rain-on-terrain ensemble: side L log-uniform in
    [lmin, lmax] rounded to 16 cells, bed amplitude U[10, 40] m, rain U[10, 100] mm/h,
    Manning n U[0.02, 0.05]."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        L = int(round(math.exp(rng.uniform(math.log(lmin), math.log(lmax))) / BW)) * BW
        out.append(dict(n=max(L, BW), amp=float(rng.uniform(10, 40)), rain=float(rng.uniform(10, 100)),
                        manning=float(rng.uniform(0.02, 0.05)), seed=seed * 1000 + k))
    return out


def ensemble_member(spec, precision="fp32", dx=5.0):
    return rain_terrain(n=spec["n"], dx=dx, amp=spec["amp"], rain=spec["rain"], manning=spec["manning"],
                        seed=spec["seed"], precision=precision)


def spherical_basin(nbx=4, nby=4, dx=0.01, lat0=-0.32, lon0=100.0, depth=10.0, amp=1.0, radius=0.05,
                    lat_c=0.0, precision="fp64", aoibnd=0, Radius=6371220.0, maxlevel=0, refine=None):
    """A lat/lon grid (dx in degrees) with a Gaussian hump at latitude lat_c over a flat bed: the
    spherical metric terms (cm, fm) enter every face and the tendencies.  maxlevel > 0 refines
    towards the hump."""
    from . import mesh_from_leaves
    if maxlevel > 0:
        if refine is None:
            def refine(l, x, y):
                s = BW * dx / (1 << l)
                return math.hypot((x + 0.5) * s - nbx * BW * dx / 2, lat0 + (y + 0.5) * s - lat_c) < (
                    0.3 if l == 0 else 0.18) * nbx * BW * dx
        leaves = quadtree_leaves(nbx, nby, maxlevel, refine)
        blocks = mesh_from_leaves(leaves, nbx, nby, dx)
        nblk = len(leaves)
        levels = [b[0] for b in leaves]
    else:
        blocks = mesh_uniform(nbx, nby, dx, 0, precision == "fp64")
        nblk = nbx * nby
        levels = [0]
    delta = dx * Radius * 3.14159265 / 180.0  # src/ReadInput.cu:1303 with pi of src/General.h:6
    p = default_param(dx=dx, delta=delta, xo=lon0, yo=lat0, nblk=nblk, nblkmem=nblkmem_for(nblk),
                      precision=precision, minlevel=min(levels), maxlevel=max(levels), aoibnd=aoibnd,
                      frictionmodel=0, cf=0.0, spherical=1, Radius=Radius)
    T = np.float64 if precision == "fp64" else np.float32
    S = p["nblkmem"] * BS
    X, Y = cell_coords_levels(blocks, p["nblkmem"], lon0, lat0, dx, T)
    zb = np.full(S, T(-depth), T)
    xc = T(lon0 + nbx * BW * dx / 2)
    r2 = (X - xc) * (X - xc) + (Y - T(lat_c)) * (Y - T(lat_c))
    h = (np.maximum(T(0.0), zb) + T(amp) * np.exp(T(-1.0) * r2 / (T(2.0) * T(radius) * T(radius))) - zb).astype(T)
    inter = np.zeros((p["nblkmem"], MW, MW), bool)
    inter[:nblk, 1:MW - 1, 1:MW - 1] = True
    h[~inter.reshape(-1)] = 0
    z = np.zeros(S, T)
    return dict(param=p, blocks=blocks, nbx=nbx, nby=nby, zb=zb, h=h, u=z.copy(), v=z.copy(), cf=z.copy())


def tsunami(nbx=32, nby=16, dx=40.0, maxlevel=2, amp=2.0, seed=5, precision="fp32", aoibnd=0, x0_frac=0.3,
            refine_depth=5.0, refine_slope=1e-4):
    """AMR tsunami: a 20 x 10 km domain (nbx x nby level-0 blocks of 16 cells of dx), a linear shelf
    from -4000 m to a +20 m beach along x, a conical island (base radius 2 km, 50 m high) at a seeded
    position, and a solitary wave of amplitude amp travelling towards the beach.  Blocks are refined
    (up to maxlevel, so dx/2, dx/4) where any cell has |grad zs| > refine_slope or h < refine_depth."""
    from . import mesh_from_leaves
    rng = np.random.default_rng(seed)
    Lx, Ly = nbx * BW * dx, nby * BW * dx
    xi, yi = float(rng.uniform(0.45, 0.7) * Lx), float(rng.uniform(0.3, 0.7) * Ly)
    x0 = x0_frac * Lx
    d0 = 4000.0 - 4020.0 * x0_frac  # still-water depth under the crest
    k = math.sqrt(3.0 * amp / (4.0 * d0 ** 3))

    def bed(x, y):
        shelf = -4000.0 + 4020.0 * x / Lx
        cone = 50.0 - 50.0 * np.hypot(x - xi, y - yi) / 2000.0  # above the shelf within 2 km
        return np.maximum(shelf, cone)

    def eta(x):
        return amp / np.cosh(k * (x - x0)) ** 2

    def refine(l, bx, by):
        s = dx / (1 << l)
        xs = (bx * BW + np.arange(-1, BW + 1) + 0.5) * s
        ys = (by * BW + np.arange(-1, BW + 1) + 0.5) * s
        X, Y = np.meshgrid(xs, ys)
        zb = bed(X, Y)
        zs = np.maximum(eta(X), zb)
        h = zs - zb
        gx = np.abs(np.diff(zs, axis=1)) / s
        gy = np.abs(np.diff(zs, axis=0)) / s
        return bool((h[1:-1, 1:-1] < refine_depth).any() or (gx > refine_slope).any() or (gy > refine_slope).any())

    leaves = quadtree_leaves(nbx, nby, maxlevel, refine)
    blocks = mesh_from_leaves(leaves, nbx, nby, dx)
    nblk = len(leaves)
    levels = [b[0] for b in leaves]
    p = default_param(dx=dx, delta=dx, nblk=nblk, nblkmem=nblkmem_for(nblk), precision=precision,
                      minlevel=min(levels), maxlevel=max(levels), aoibnd=aoibnd, frictionmodel=-1, cf=0.025)
    T = np.float64 if precision == "fp64" else np.float32
    S = p["nblkmem"] * BS
    X, Y = cell_coords_levels(blocks, p["nblkmem"], 0.0, 0.0, dx, np.float64)
    zb64 = bed(X, Y)
    e = eta(X)
    h64 = np.maximum(np.maximum(e, zb64) - zb64, 0.0)
    u64 = np.where(h64 > 1.0, e * np.sqrt(9.81 / np.maximum(h64, 1.0)), 0.0)
    inter = np.zeros((p["nblkmem"], MW, MW), bool)
    inter[:nblk, 1:MW - 1, 1:MW - 1] = True
    m = inter.reshape(-1)
    zb = np.where(m, zb64, 0.0).astype(T)
    h = np.where(m, h64, 0.0).astype(T)
    u = np.where(m, u64, 0.0).astype(T)
    z = np.zeros(S, T)
    # the bed as a raster at the finest cell size, for re-sampling after in-run adaptation
    rdx = dx / (1 << maxlevel)
    gx = np.arange(0.0, Lx + 0.5 * rdx, rdx)
    gy = np.arange(0.0, Ly + 0.5 * rdx, rdx)
    Xg, Yg = np.meshgrid(gx, gy)
    raster = (bed(Xg, Yg).astype(np.float32), 0.0, 0.0, rdx)
    return dict(param=p, blocks=blocks, nbx=nbx, nby=nby, zb=zb, h=h, u=u, v=z.copy(), cf=np.full(S, T(0.025), T),
                leaves=leaves, bed_raster=raster, maxlevel=maxlevel,
                criteria=[("zs", 2, refine_slope, 0.0), ("h", 1, -1.0, refine_depth)])


def run_adaptive(domain, loop, nsteps, nbx, nby, criteria, every=10, coarsen=True, endtime=1e300, log=None):
    """Runs nsteps on a multi-level Domain, re-adapting the mesh every `every` steps from the criteria
    on the GPU state (Domain.adapt).  Returns per-adaptation records (step, blocks before, after,
    refined, coarsened, seconds spent adapting)."""
    import time
    recs = []
    done = 0
    while done < nsteps:
        n = min(every, nsteps - done)
        domain.run(loop, n, endtime=endtime)
        done += n
        if done >= nsteps or loop.totaltime >= endtime:
            break
        t0 = time.perf_counter()
        n0, n1, nref, nco = domain.adapt(criteria, nbx, nby, coarsen=coarsen)
        recs.append((done, n0, n1, nref, nco, time.perf_counter() - t0))
        if log:
            log("adapt at step %d: %d -> %d blocks (%d refined, %d coarsened)" % (done, n0, n1, nref, nco))
    return recs


def block_any(blocks, nblk, nblkmem, dx, fn, precision="fp64"):
    """Per block (nblkmem entries): whether fn(x, y) holds at any interior cell centre."""
    T = np.float64 if precision == "fp64" else np.float32
    X, Y = cell_coords_levels(blocks, nblkmem, 0.0, 0.0, dx, T)
    hit = np.asarray(fn(X.astype(np.float64), Y.astype(np.float64)), bool).reshape(nblkmem, MW, MW)
    inter = np.zeros((nblkmem, MW, MW), bool)
    inter[np.asarray(blocks["active"][:nblk]), 1:MW - 1, 1:MW - 1] = True
    return (hit & inter).any(axis=(1, 2))


def adapt_mesh(nbx, nby, dx, maxlevel, fn, precision="fp64", maxiter=None):
    """The reference's set-up adaptation loop (src/Adaptation.cu:17-121): starting from the uniform
    level-0 mesh, refine every block where fn(x, y) holds at some cell, with the reference's sanity
    cascade and block numbering (bgf_adapt_refine), until no block is added.  Returns
    (blocks, nblk, nblkmem)."""
    from . import adapt_refine
    blocks = mesh_uniform(nbx, nby, dx, 0, precision == "fp64")
    nblk = nbx * nby
    nblkmem = nblkmem_for(nblk)
    for _ in range(maxiter or maxlevel + 1):
        p = default_param(dx=dx, delta=dx, nblk=nblk, nblkmem=nblkmem, maxlevel=maxlevel, precision=precision)
        flags = block_any(blocks, nblk, nblkmem, dx, fn, precision)
        nb2, n2, m2 = adapt_refine(p, blocks, flags, nbx, nby)
        if n2 == nblk:
            break
        blocks, nblk, nblkmem = nb2, n2, m2
    return blocks, nblk, nblkmem
