"""Named workloads of the erosion hot path (inputs and parameters).

``reference``  the reference demo: 512x512 bounded map, reference defaults,
               4096 particles per World::erode call
               (SimpleWindErosion.cpp:217, cellpool.h:184-198).
``cpu256``     256x256 periodic map, fBm height + 0.05 sediment layer,
               wind normalize(1,0,1), suspension 1e-4, abrasion 1e-4,
               roughness (maxdiff) 0.005, settling 0.01; 2,000 particles x
               100 cycles.
``gpu1024``    1024x1024 periodic map, same recipe plus 8 square pyramids
               (height 0.3-0.6, base 40-120 cells) in the height layer;
               1,000,000 particles x 200 cycles.

All three use the step-synchronous window (one fly() call per snapshot
window): particles in this dynamics crowd into lee-side cells (tens per cell
per step at one particle per cell), and longer windows let those cells'
cascades overshoot.

Surface height = fBm (+ pyramids) + sediment thickness: the library's
height layer is the surface the reference's cell->height describes
(cellpool.h:228), the sediment layer the loose part on top of bedrock.
"""
import numpy as np

from .world import World, Wind, generate_terrain, glm_normalize

WORKLOADS = {
    "reference": dict(nx=512, ny=512, particles=4096, cycles=1, periodic=False, layers=1, window=1,
                      params=dict(), terrain=None, sediment=None),
    "cpu256": dict(nx=256, ny=256, particles=2000, cycles=100, periodic=True, layers=2, window=1,
                   params=dict(suspension=1e-4, abrasion=1e-4, maxdiff=0.005, settling=0.01,
                               pspeed=glm_normalize((1.0, 0.0, 1.0))),
                   terrain=dict(octaves=6, base_cycles=4.0, n_pyramids=0), sediment=0.05),
    "gpu1024": dict(nx=1024, ny=1024, particles=1_000_000, cycles=200, periodic=True, layers=2, window=1,
                    params=dict(suspension=1e-4, abrasion=1e-4, maxdiff=0.005, settling=0.01,
                                pspeed=glm_normalize((1.0, 0.0, 1.0))),
                    terrain=dict(octaves=6, base_cycles=4.0, n_pyramids=8), sediment=0.05),
    # the surrogate's training maps and its 1024^2 rollout map (the fBm
    # recipe with the 8 obstacles); two layers, one particle per cell, mass
    # kept on the map (deposit on death).  The 256^2 training maps have the
    # rollout map's features per cell: one fBm period per 256 cells and, on
    # every other map, one obstacle (0.5 per 256^2 of the rollout map).
    "surr256": dict(nx=256, ny=256, particles=65536, cycles=50, periodic=True, layers=2, window=1,
                    params=dict(suspension=1e-4, abrasion=1e-4, maxdiff=0.005, settling=0.01,
                                pspeed=glm_normalize((1.0, 0.0, 1.0))),
                    terrain=dict(octaves=6, base_cycles=1.0, n_pyramids=1), sediment=0.05),
    "surr1024": dict(nx=1024, ny=1024, particles=1_000_000, cycles=50, periodic=True, layers=2, window=1,
                     params=dict(suspension=1e-4, abrasion=1e-4, maxdiff=0.005, settling=0.01,
                                 pspeed=glm_normalize((1.0, 0.0, 1.0))),
                     terrain=dict(octaves=6, base_cycles=4.0, n_pyramids=8), sediment=0.05),
}


def make_inputs(name, seed=1, nx=None, ny=None, n_pyramids=None):
    """Surface height and sediment arrays (float32, [nx, ny]) of a workload
    (n_pyramids overrides the workload's obstacle count)."""
    import numpy as np

    w = WORKLOADS[name]
    nx = nx or w["nx"]
    ny = ny or w["ny"]
    if w["terrain"] is None:
        # cellpool.h:380-385,416-428: Gaussian bump at (W/4, H/2), scaled to [0, 0.5]
        x = np.arange(nx, dtype=np.float32)[:, None] / np.float32(nx)
        y = np.arange(ny, dtype=np.float32)[None, :] / np.float32(ny)
        d2 = (x - np.float32(0.25)) ** 2 + (y - np.float32(0.5)) ** 2
        h = np.exp(-d2 * np.float32(nx * 0.2)).astype(np.float32)
        lo = min(0.0, float(h.min()))
        h = (0.5 * (h - lo) / (float(h.max()) - lo)).astype(np.float32)
        return h, None
    t = w["terrain"]
    base = generate_terrain(nx, ny, seed, octaves=t["octaves"], base_cycles=t["base_cycles"],
                            n_pyramids=t["n_pyramids"] if n_pyramids is None else n_pyramids)
    if w["sediment"] is None:
        return base.astype(np.float32), None
    sed = np.full((nx, ny), np.float32(w["sediment"]), np.float32)
    return (base + sed).astype(np.float32), sed


def apply(name, seed=1, **over):
    """Set the Wind/World statics of a workload; return the per-map World options."""
    w = WORKLOADS[name]
    p = dict(w["params"])
    p.update(over.pop("params", {}))
    for k in ("suspension", "gravity", "boundarylayer", "pspeed"):
        if k in p:
            setattr(Wind, k, p.pop(k))
    for k in ("lrate", "maxdiff", "settling"):
        if k in p:
            setattr(World, k, p.pop(k))
    World.SEED = seed
    opts = dict(periodic=w["periodic"], layers=w["layers"], window=w["window"], abrasion=p.pop("abrasion", 0.0))
    opts.update(over)
    return opts


def params(name, seed=1, **over):
    """swe_params of a workload (sets the statics; no device needed)."""
    from .world import build_params

    return build_params(**apply(name, seed, **over))


def make_world(name, seed=1, nx=None, ny=None, device=0, **over):
    """A World loaded with the workload's inputs and parameters."""
    w = WORKLOADS[name]
    nx = nx or w["nx"]
    ny = ny or w["ny"]
    opts = apply(name, seed, **over)
    world = World(nx, ny, device=device, **opts)
    h, s = make_inputs(name, seed, nx, ny)
    world.upload(h, s)
    return world


BATCH_SIDES = (128, 192, 256, 384, 512, 768, 1024)


def batch64(seed=1, n=64, sides=BATCH_SIDES, elevation_deg=45.0):
    """The multi-terrain workload: `n` terrains whose sides are drawn
    log-uniformly from `sides` (rectangles allowed), each with its own seed
    and a prevailing wind of uniform azimuth at `elevation_deg` above the
    horizon: pspeed = (cos e cos a, sin e, cos e sin a); particles per cycle =
    area.  Returns a list of dicts (nx, ny, seed, pspeed, particles)."""
    import math

    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    e = math.radians(elevation_deg)
    for k in range(n):
        nx = int(sides[rng.integers(len(sides))])
        ny = int(sides[rng.integers(len(sides))])
        az = float(rng.uniform(0.0, 2.0 * math.pi))
        ps = glm_normalize((math.cos(e) * math.cos(az), math.sin(e), math.cos(e) * math.sin(az)))
        out.append(dict(nx=nx, ny=ny, seed=int(rng.integers(1, 2 ** 31)), pspeed=ps, particles=nx * ny))
    return out


def make_batch(terrains, device=0, layers=2, sediment=0.05, **opts):
    """A Batch loaded with fBm terrains (+ a uniform sediment layer) and
    the cpu256 workload's rates."""
    from .world import Batch

    for k, v in WORKLOADS["cpu256"]["params"].items():
        if k in ("suspension", "gravity", "boundarylayer"):
            setattr(Wind, k, v)
        elif k in ("lrate", "maxdiff", "settling"):
            setattr(World, k, v)
    b = Batch([(t["nx"], t["ny"]) for t in terrains], device=device, layers=layers,
              abrasion=WORKLOADS["cpu256"]["params"]["abrasion"], **opts)
    for k, t in enumerate(terrains):
        b.set_terrain(k, t["seed"], t["pspeed"])
        h = generate_terrain(t["nx"], t["ny"], t["seed"])
        s = np.full((t["nx"], t["ny"]), np.float32(sediment), np.float32) if layers == 2 else None
        b.upload(k, (h + (s if s is not None else 0)).astype(np.float32), s)
    return b
