"""Writes the golden vectors under tests/golden/ from the CPU oracle.

The reference is Julia and cannot run here, so the fixtures are produced by
the oracle, whose functions are pinned by the reference's closed-form tests
(tests/test_oracle_*.py).  They freeze seeded inputs and float64 outputs so
the GPU engine can be checked without re-running the oracle, and so the same
inputs can be loaded into the reference on a machine with Julia.

Run from the repository root:  python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402

SPECS = {
    # name: (resolution, domain dims, lower, upper, processes, dt)
    "les_12x14x16": ((12, 14, 16), (2 * np.pi, 4 / 3 * np.pi, 1.0), ("rough", 5e-3), ("freeslip",),
                     lambda: O.incompressible_flow(1e-4, constant_forcing=(1, 0.2),
                                                   sgs_model=O.StaticSmagorinskyModel()), 1e-3),
    "channel_8x8x6": ((8, 8, 6), (2 * np.pi, np.pi, 2.0), ("smooth",), ("smooth",),
                      lambda: O.incompressible_flow(1 / 180, constant_forcing=1), 1e-3),
    "pressure_4x4x16": ((4, 4, 16), (2 * np.pi, 2 * np.pi, 1.0), ("smooth",), ("smooth",),
                        lambda: [O.Pressure()], 0.0),
}


def boundary(spec):
    kind = spec[0]
    if kind == "rough":
        return O.RoughWall(spec[1])
    if kind == "freeslip":
        return O.FreeSlipBoundary()
    return O.SmoothWall()


def build(name):
    res, dims, lo, hi, procs, dt = SPECS[name]
    return O.Model(res, O.Domain(dims, boundary(lo), boundary(hi)), procs()), dt


def seeded_state(model, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    K1, K2 = model.grid.fdsize()
    k1, k2 = O.wavenumbers(model.grid)
    scale = np.exp(-0.4 * np.sqrt(k1[:, None] ** 2 + k2[None, :] ** 2))[:, :, None]
    out = {}
    km = model.grid.k2max
    for f in model.fields:
        n = model.state[f].shape[2]
        a = (rng.standard_normal((K1, K2, n)) + 1j * rng.standard_normal((K1, K2, n))) * 0.1 * scale
        a[0, K2 - km:, :] = np.conj(a[0, 1:km + 1, :][::-1, :])
        a[0, 0, :] = a[0, 0, :].real
        out[f] = a
    if "vel1" in out:
        out["vel1"][0, 0, :] += 1.0
    return out


def main():
    for i, name in enumerate(SPECS):
        m, dt = build(name)
        st = seeded_state(m, 100 + i)
        data = {f"state_{f}": v for f, v in st.items()}
        for f in m.fields:
            m.state[f][...] = st[f]
        rates = {f: np.zeros_like(a) for f, a in m.state.items()}
        O.compute_rates(m, rates, m.state)
        data.update({f"rates_{f}": v for f, v in rates.items()})
        O.apply_projections(m, m.state)
        data.update({f"projected_{f}": v.copy() for f, v in m.state.items()})
        if dt > 0:
            for f in m.fields:
                m.state[f][...] = st[f]
            O.evolve(m, dt, dt)
            data.update({f"step_{f}": v.copy() for f, v in m.state.items()})
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **data)
        print("wrote", name)


if __name__ == "__main__":
    main()
