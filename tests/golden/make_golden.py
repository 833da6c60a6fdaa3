"""Writes the golden vectors of tests/golden from the reference build.

Runs oracle/_ref/ref_driver (the reference's own FFT-free Fortran routines,
compiled from their sources by oracle/Makefile) on seeded synthetic inputs and
stores inputs and outputs as .npz fixtures.  Run in the container that holds
the reference sources:  make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import Oracle, run_reference  # noqa: E402
from simox_amd.systems import config, tiny, tiny_alloy  # noqa: E402


T_BEREND = 600.0      # K, target of VV_NVT2_BEREND
ALPHA_LANGEVIN = 0.1  # friction of VV_NVT1/2_STOCH (run at T_given = 0)


def one(name, s, seed, with_dudq=True):
    rng = np.random.default_rng(seed)
    q = rng.normal(size=s.n) * 0.3
    q[s.ids == 4] = -np.abs(q[s.ids == 4])
    f0 = rng.normal(size=(s.n, 3))
    dud = None
    o = Oracle(s.n, s.box, s.fft, s.rcut)
    o.cell_sort(s.x)
    if with_dudq:
        nlist, lst, asym, sym = o.ctip_matrix(s.ids, s.x)
        phi = rng.normal(size=s.n)
        dud = (nlist, lst, asym, sym, phi)
    # VV_NVE1_REFLECT case: z shifted so that the 5 highest atoms sit on or
    # above the reflecting wall at box_z / 2
    zshift = 0.5 * s.box[2] - np.sort(s.x[:, 2])[-5]
    ref = run_reference(s.ids, s.x, s.v, f0, q, s.box, s.fft, s.rcut, s.dt, dudq_inputs=dud,
                        t_given=T_BEREND, alpha=ALPHA_LANGEVIN, zshift=zshift)
    out = dict(ids=s.ids, x=s.x, v=s.v, f0=f0, q=q, box=s.box, fft=np.array(s.fft),
               rcut=np.array(s.rcut), dt=np.array(s.dt),
               ref_npt=ref["npt"], ref_pair=ref["pair"], ref_f=ref["f"],
               ref_E=np.array([ref["Epair"], ref["Edir"], ref["Ectip"], ref["Egroup"], ref["Eself"]]),
               ref_x1=ref["x1"], ref_v1=ref["v1"], ref_v2=ref["v2"], ref_mv2=np.array(ref["mv2"]),
               ref_rdf=ref["rdf"],
               t_given=np.array(T_BEREND), alpha=np.array(ALPHA_LANGEVIN),
               ref_vb=ref["vb"], ref_mv2b=np.array(ref["mv2b"]), ref_xs=ref["xs"],
               ref_vs=ref["vs"], ref_vs2=ref["vs2"], ref_fs2=ref["fs2"],
               ref_mv2s=np.array(ref["mv2s"]), zshift=np.array(zshift),
               ref_xr=ref["xr"], ref_vr=ref["vr"], ref_xr2=ref["xr2"], ref_vr2=ref["vr2"])
    # links only for occupied slots (compact)
    out["ref_link"] = np.concatenate([ref["link"][c, :ref["npt"][c]] for c in range(len(ref["npt"]))])
    if s.n <= 400:
        out["ref_psi"] = ref["psi"]
    if with_dudq:
        out["dudq_phi"] = dud[4]
        out["ref_g"] = ref["g"]
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path))


def oxidation(name, s, seed):
    """OXIDE_CHECK (util.f90:152-188) and NEW_O2 (datafmt.f90:98-139) of the
    reference build: per record, the new-dimer flag for each q_crit, then the
    dimers NEW_O2 places after RANDOM_SEED(put = (seed, 1, 1, ...)).  Records:
    the slab as built (its last O above the top Al), the same with those O
    moved 3 A below the top Al, and a record whose last 2 n_o2 atoms include
    Al (OXIDE_CHECK skips them)."""
    rng = np.random.default_rng(seed)
    q = -np.abs(rng.normal(size=s.n) * 0.3)
    q_crits = [-5.0, -0.2, -0.1, 0.0, 5.0]
    top_al = s.x[s.ids == 1, 2].max()
    x_low = s.x.copy()
    x_low[-4:, 2] = top_al - 3.0
    ids_mix = s.ids.copy()
    ids_mix[-1] = 1
    recs = [(s.ids, s.x, 2, 7, 10.0), (s.ids, x_low, 2, 8, 11.5), (ids_mix, s.x, 3, 9, 12.0),
            (s.ids, s.x, 1, 10, 9.0)]
    out = dict(box=s.box, q=q, t_given=np.array(T_BEREND), q_crit=np.array(q_crits),
               nrec=np.array(len(recs)))
    for k, (ids, x, n_o2, sd, z_new) in enumerate(recs):
        ref = run_reference(ids, x, s.v, np.zeros((s.n, 3)), q, s.box, s.fft, s.rcut, s.dt,
                            t_given=T_BEREND, oxid=(n_o2, sd, z_new, q_crits))
        out.update({"ids%d" % k: ids, "x%d" % k: x, "n_o2_%d" % k: np.array(n_o2),
                    "seed%d" % k: np.array(sd), "z_new%d" % k: np.array(z_new),
                    "ref_new_o2_%d" % k: ref["new_o2"], "ref_o2_x%d" % k: ref["o2_x"],
                    "ref_o2_v%d" % k: ref["o2_v"], "ref_o2_id%d" % k: ref["o2_id"],
                    "ref_o2_q%d" % k: ref["o2_q"]})
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path))


if __name__ == "__main__":
    one("tiny_s1", tiny(seed=1), 11)
    one("tiny_s2", tiny(seed=2, nx=3, ny=3, nz=4, n_o=8, fft=(12, 12, 24), rcut=5.0), 12)
    one("c1_s1", config("c1", seed=1), 21)
    # Al slab with Ni and Fe substitutions: the AL_NI and METAL_HETERO branches
    one("alloy_s1", tiny_alloy(seed=1), 31)
    oxidation("oxid_s1", tiny(seed=1), 41)
