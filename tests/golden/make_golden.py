"""Writes tests/golden/kats.json: known-answer vectors for the particle path.

Every value here comes from a closed form, from a brute-force enumeration, or
from a number printed by the reference itself (its run logs and unit-test
constants). Nothing is computed with the CPU checker under oracle/, so the
fixtures can pin it.

Sources (SueHeir/MDDEM v0.1.3):
  * examples/granular_gas_benchmark/last_run.log:13-14,18 -- pair/ghost cutoff,
    13 forward stencil cells, dt = 0.15 x Rayleigh time
  * crates/dem_granular/src/contact.rs:2316-2360 -- Hertz analytic force
  * crates/dem_granular/src/contact.rs:2215-2265 -- delta^(3/2) scaling
  * crates/dem_atom/src/lib.rs:679-764 -- beta, friction, E* closed forms
  * crates/md_lj/src/lib.rs:686-752 -- LJ two-body sign tests
  * crates/mddem_core/src/neighbor.rs:1164-1266 -- 30-atom trig-curve pair set
  * crates/md_lattice/src/lib.rs:517-560 -- FCC count / density / zero COM
  * examples/bench_hertz_rebound/validate.py:12-13,59-64 -- contact time and
    peak overlap of a Hertz impact

Run: python tests/golden/make_golden.py
"""
import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    out = {}
    E, nu, e, mu, rho = 8.7e9, 0.3, 0.95, 0.4, 2500.0

    # material mixing for one material with itself
    log_e = math.log(e)
    out["glass_beta"] = -log_e / math.sqrt(math.pi ** 2 + log_e ** 2)
    out["glass_friction"] = mu
    out["glass_e_eff"] = E / (2.0 * (1.0 - nu * nu))
    out["glass_g_eff"] = E / (4.0 * (2.0 - nu) * (1.0 + nu))
    # two materials (dem_atom lib.rs:705-750): glass + steel
    E2, nu2, e2, mu2 = 200e9, 0.28, 0.8, 0.3
    le = math.log(math.sqrt(e * e2))
    out["mix_beta"] = -le / math.sqrt(math.pi ** 2 + le ** 2)
    out["mix_friction"] = math.sqrt(mu * mu2)
    out["mix_e_eff"] = 1.0 / ((1 - nu * nu) / E + (1 - nu2 * nu2) / E2)

    # Rayleigh time step printed by the reference (last_run.log:18)
    out["granular_gas_dt"] = 4.400938879498286e-7
    out["granular_gas_dt_factor"] = 0.15
    out["granular_gas_pair_cutoff"] = 0.0022
    out["granular_gas_ghost_cutoff"] = 0.0024
    out["granular_gas_forward_stencil"] = 13

    # Hertz normal force for two glass spheres, r = 1 mm, delta = 5e-5 m
    r = 1e-3
    r_eff = r / 2.0
    delta = 5e-5
    out["hertz_radius"] = r
    out["hertz_delta"] = delta
    out["hertz_force"] = (4.0 / 3.0) * out["glass_e_eff"] * math.sqrt(r_eff) * delta ** 1.5
    out["hertz_scaling_deltas"] = [1e-5, 2e-5, 4e-5, 6e-5, 8e-5]

    # Hertz impact closed forms (Johnson 1985) for a head-on glass pair
    m = rho * 4.0 / 3.0 * math.pi * r ** 3
    m_eff = m / 2.0
    v0 = 0.5
    e_star = out["glass_e_eff"]
    out["impact_v0"] = v0
    out["impact_contact_time"] = 2.87 * (m_eff ** 2 / (r_eff * e_star ** 2 * v0)) ** 0.2
    out["impact_peak_overlap"] = (15.0 * m_eff * v0 ** 2 / (16.0 * r_eff ** 0.5 * e_star)) ** 0.4

    # LJ 12-6 pair force along x for eps = sigma = 1: F_x on the atom at +r
    lj = {}
    for rr in (0.9, 1.0, 1.12, 1.2, 1.5, 2.0, 2.4, 3.0):
        f = 24.0 * (2.0 * rr ** -13 - rr ** -7) if rr < 2.5 else 0.0
        lj[repr(rr)] = f
    out["lj_force_on_plus_x"] = lj
    out["lj_zero_force_r"] = 2.0 ** (1.0 / 6.0)
    out["lj_coeffs"] = [48.0, 24.0, 6.25]

    # 30 atoms on a trig curve, radius 0.3, skin 1.0, box [0,4]^3 non periodic
    n = 30
    pos = []
    for i in range(n):
        pos.append([math.sin(i * 0.31) * 1.5 + 2.0,
                    math.cos(i * 0.67) * 1.5 + 2.0,
                    math.sin(i * 0.97) * 1.5 + 2.0])
    pairs = []
    for i in range(n):
        for j in range(i + 1, n):
            d = math.sqrt(sum((pos[j][k] - pos[i][k]) ** 2 for k in range(3)))
            if d < 0.6:
                pairs.append([i, j])
    out["trig30_pos"] = pos
    out["trig30_pairs"] = pairs

    # FCC shells within 2.8 at rho = 0.85: 12 + 6 + 24 + 12 + 24 = 78
    a = (4.0 / 0.85) ** (1.0 / 3.0)
    shells = {}
    rng = range(-3, 4)
    basis = [(0, 0, 0), (0.5, 0.5, 0), (0.5, 0, 0.5), (0, 0.5, 0.5)]
    cnt = 0
    for ix in rng:
        for iy in rng:
            for iz in rng:
                for b in basis:
                    x = (ix + b[0]) * a
                    y = (iy + b[1]) * a
                    z = (iz + b[2]) * a
                    d2 = x * x + y * y + z * z
                    if 0 < d2 < 2.8 ** 2:
                        cnt += 1
                        key = round(math.sqrt(d2) / a, 6)
                        shells[key] = shells.get(key, 0) + 1
    out["fcc_neighbors_within_2p8"] = cnt
    out["fcc_shells"] = {repr(k): v for k, v in sorted(shells.items())}

    with open(os.path.join(HERE, "kats.json"), "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print("wrote", os.path.join(HERE, "kats.json"))


if __name__ == "__main__":
    main()
