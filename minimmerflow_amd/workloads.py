"""Seeded case sets of the batched mode.

mixed_2d_cases() draws independent 2D cases: the mesh side uniformly from
{32, 48, ..., 256} (s = 32 + 16 floor(15 u)); the first half of the cases
are isentropic vortices (problem.cpp:247-318) with beta ~ U[2,6] and centre
~ U[-2,2]^2 on [-5,5]^2 with periodic borders (the exact solution wrapped);
the second half radial Sod (problem.cpp:319-355) on [-1,1]^2 with r0 = 0.4,
inside rho = p = 1, outside rho = 1 / r_rho and p = 1 / r_p with r_rho, r_p ~
U[4,16], transmissive (free-flow) borders.  The draws come from SplitMix64
with u = (x >> 11) 2^-53, in case order (side first), so a seed fixes the
set on every machine.
"""
import numpy as np

MASK64 = (1 << 64) - 1
DEFAULT_SEED = 0x6D6D66
SIDES = tuple(range(32, 257, 16))


class SplitMix64:
    """Steele, Lea and Flood's SplitMix64 generator; uniform() takes the top 53 bits."""

    def __init__(self, seed):
        self.state = seed & MASK64

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / (1 << 53)))

    def choice(self, seq):
        return seq[min(int(self.uniform() * len(seq)), len(seq) - 1)]


def mixed_2d_case_params(ncases=512, seed=DEFAULT_SEED):
    """Plain parameters of each case (no library needed)."""
    rng = SplitMix64(seed)
    out = []
    for i in range(ncases):
        side = rng.choice(SIDES)
        if i < ncases // 2:
            beta = rng.uniform(2.0, 6.0)
            cx, cy = rng.uniform(-2.0, 2.0), rng.uniform(-2.0, 2.0)
            out.append({"kind": "vortex_xy", "side": side, "beta": beta, "center": (cx, cy)})
        else:
            dr = rng.uniform(4.0, 16.0)
            pr = rng.uniform(4.0, 16.0)
            out.append({"kind": "radsod", "side": side, "density_ratio": dr, "pressure_ratio": pr})
    return out


def problem_of(mmf, prm):
    if prm["kind"] == "vortex_xy":
        c = prm["center"]
        return mmf.problem("vortex_xy", 2, vortex_beta=prm["beta"], vortex_center=[c[0], c[1], 0.0],
                           vortex_periodic=1)
    # rho = p / T with R = 1: inside rho = p = 1 (T = 1); outside p = 1/pr, rho = 1/dr
    dr, pr = prm["density_ratio"], prm["pressure_ratio"]
    return mmf.problem("radsod", 2, domain_origin=[-1.0, -1.0, -1.0], domain_length=2.0, sod_r0=0.4,
                       sod_p_in=1.0, sod_t_in=1.0, sod_p_out=1.0 / pr, sod_t_out=dr / pr)


def side_bc_of(mmf, prm):
    """Periodic vortices, transmissive radial Sod."""
    return [mmf.BC_PERIODIC] * 4 if prm["kind"] == "vortex_xy" else [mmf.BC_FREE_FLOW] * 4


def mixed_2d_cases(mmf, ncases=512, seed=DEFAULT_SEED):
    """(grid descriptors, problems, parameters) of the seeded mixed case set."""
    params = mixed_2d_case_params(ncases, seed)
    problems = [problem_of(mmf, p) for p in params]
    descs = []
    for pb, p in zip(problems, params):
        d = mmf.problem_grid_desc(pb, p["side"])
        for s, b in enumerate(side_bc_of(mmf, p)):
            d.side_bc[s] = b
        descs.append(d)
    return descs, problems, params


def side_histogram(params):
    return {s: sum(1 for p in params if p["side"] == s) for s in SIDES}


def total_cells(params):
    return int(np.sum([p["side"] ** 2 for p in params]))
