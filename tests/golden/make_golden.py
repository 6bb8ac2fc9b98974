"""Write the Mode R ("reference") golden vectors from the reference implementation.

The reference's fast sweep ``LatticeQMC._update_step`` (lqmc/lqmc.py:301-347)
follows no invariant (its update coefficients and wrap are not the textbook
ones), so only running it pins a restatement. This script is the one place in
the repository that imports the reference; it runs in the development
container only (the reference checkout is at /root/reference there) and writes
small ``.npz`` fixtures next to itself. Nothing else reads the reference.

Recorded per case (all float64 unless noted):
  K            kinetic matrix of the model (input of the sweep, lqmc/hubbard.py:91-111)
  beta, L, u, t, mu, seed
  field0       int8 (N, L) HS field after construction (configuration.py:120-124)
  uniforms     (n_sweeps, L*N) the np.random.rand() values the sweeps consumed
  ratio, acc   (n_sweeps, L*N) per-attempt ratio and accept flag (lqmc.py:313-317)
  field        int8 (n_sweeps, N, L) field after each sweep
  g_up, g_dn   (n_sweeps, N, N) G returned by each _update_step call
  meas_g       (2, N, N) measure_loop result over the recorded sweeps (lqmc.py:356-375)

Run:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def _import_reference():
    os.environ.setdefault("MPLBACKEND", "Agg")
    sys.path.insert(0, REF)
    import lqmc  # noqa: E402
    from lqmc.lqmc import LatticeQMC  # noqa: E402
    return lqmc, LatticeQMC


def record_mode_r(size, u, beta, L, seed, n_sweeps, mu=0.0):
    lqmc, LatticeQMC = _import_reference()

    class Recorder(LatticeQMC):
        def __init__(self, *a, **k):
            self.rec = []
            super().__init__(*a, **k)

        def _debug(self, i, l):  # called once per attempt, after the accept decision
            self.rec.append((l, i, float(self.ratio), bool(self.acc)))

    model = lqmc.HubbardModel(u=u, t=1.0, mu=mu)
    model.build_square(size)
    np.random.seed(seed)
    solver = Recorder(model, beta, L, warmup=0, sweeps=n_sweeps, log_lvl=None)
    field0 = solver.config.config.copy()
    n = model.n_sites

    # Capture the uniform stream by replaying the global RNG state around each sweep.
    ratios, accs, fields, gups, gdns, unis = [], [], [], [], [], []
    gsum_up = np.zeros((n, n))
    gsum_dn = np.zeros((n, n))
    for _ in range(n_sweeps):
        state = np.random.get_state()
        solver.rec = []
        g_up, g_dn = solver._update_step()
        after = np.random.get_state()
        rs = np.random.RandomState()
        rs.set_state(state)
        u_sweep = rs.rand(L * n)
        chk = np.random.RandomState()
        chk.set_state(after)
        # the stream position after the sweep must equal the one after drawing L*N values
        assert np.array_equal(rs.get_state()[1], chk.get_state()[1])
        assert len(solver.rec) == L * n
        order = [(l, i) for l in reversed(range(L)) for i in range(n)]
        assert [(r[0], r[1]) for r in solver.rec] == order
        ratios.append([r[2] for r in solver.rec])
        accs.append([r[3] for r in solver.rec])
        fields.append(solver.config.config.copy())
        gups.append(g_up.copy())
        gdns.append(g_dn.copy())
        unis.append(u_sweep)
        gsum_up += g_up
        gsum_dn += g_dn
    meas = np.asarray([gsum_up, gsum_dn]) / n_sweeps
    return dict(
        K=model.ham_kinetic(), beta=beta, L=L, u=u, t=1.0, mu=mu, seed=seed,
        field0=field0.astype(np.int8), uniforms=np.asarray(unis),
        ratio=np.asarray(ratios), acc=np.asarray(accs, dtype=bool),
        field=np.asarray(fields, dtype=np.int8), g_up=np.asarray(gups),
        g_dn=np.asarray(gdns), meas_g=meas,
    )


def record_lattices():
    lqmc, _ = _import_reference()
    out = {}
    for side in (2, 3, 4, 5, 8):
        m = lqmc.HubbardModel(u=4, t=1.0, mu=0.0)
        m.build_square(side)
        out[f"K_square_{side}"] = m.ham_kinetic()
    m = lqmc.HubbardModel(u=0, t=1.0, mu=0.0)
    m.build(10, cycling=None)  # exact.py:44-46 open 10-site chain
    out["K_chain_10_open"] = m.ham_kinetic()
    m = lqmc.HubbardModel(u=2, t=1.0, mu=None)
    m.build(6)  # default cycling=0: periodic 6-site ring, mu=U/2 default
    out["K_chain_6_periodic_mu_default"] = m.ham_kinetic()
    m = lqmc.HubbardModel(u=2, t=0.5, mu=0.25)
    m.build(3, 4, cycling=(0, 1))
    out["K_rect_3x4"] = m.ham_kinetic()
    # other prefabs (lattice.py:365-421): open hexagonal (two-atom cell), cubic open and
    # periodic, and a square lattice carrying two distance shells
    from lqmc.lattice import Lattice
    for name, make, shape, pbc in (("K_hex_2x2_open", Lattice.hexagonal, (2, 2), None),
                                   ("K_hex_3x2_open", Lattice.hexagonal, (3, 2), None),
                                   ("K_cubic_3x3x3_open", Lattice.cubic, (3, 3, 3), None),
                                   ("K_cubic_3x3x3_pbc", Lattice.cubic, (3, 3, 3), (0, 1, 2)),
                                   ("K_square_3x5_open_nn2", lambda: Lattice.square(neighbour_dist=2), (3, 5), None)):
        latt = make()
        latt.build(shape)
        if pbc is not None:
            latt.set_periodic_boundary(pbc)
        m = lqmc.HubbardModel(u=4, t=1.0, mu=0.5)
        m.lattice = latt
        out[name] = m.ham_kinetic()
    return out


def main():
    if "--lattices-only" in sys.argv:
        lat = record_lattices()
        np.savez_compressed(os.path.join(HERE, "lattices.npz"), **lat)
        print("lattices.npz", list(lat))
        return
    cases = {
        "mode_r_2x2.npz": dict(size=2, u=4.0, beta=2.0, L=16, seed=0, n_sweeps=2),
        "mode_r_4x4.npz": dict(size=4, u=4.0, beta=2.0, L=16, seed=0, n_sweeps=2),
        "mode_r_4x4_u0.npz": dict(size=4, u=0.0, beta=2.0, L=16, seed=1, n_sweeps=1),
    }
    for name, kw in cases.items():
        data = record_mode_r(**kw)
        np.savez_compressed(os.path.join(HERE, name), **data)
        print(name, {k: np.shape(v) for k, v in data.items()})
    lat = record_lattices()
    np.savez_compressed(os.path.join(HERE, "lattices.npz"), **lat)
    print("lattices.npz", list(lat))


if __name__ == "__main__":
    main()
