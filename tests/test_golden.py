"""The committed golden fixtures (tests/golden/, written by
tests/golden/make_fixtures.py) on the CPU:

- the data of the reference's own helper tests (bounding_box,
  test/test_mask.py:19-31; the disk area, test/test_shape.py:31-35) against
  the package's helpers;
- the short oracle iterations of iterates_small.npz reproduced by the oracle
  (a change to the oracle that moves them fails here);
- the final-error distributions of the benchmark workloads: structure, and
  one seed of the 256^2 CPU workload recomputed.
"""
import json
import os

import numpy as np
import pytest

from als_speckle_amd import masking, synth
from oracle import phasing_c128 as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_json(name):
    with open(os.path.join(GOLD, name)) as fh:
        return json.load(fh)


def small_cases():
    z = np.load(os.path.join(GOLD, "iterates_small.npz"))
    names = sorted({k.split("/")[0] for k in z.files})
    return z, names


def test_reference_bounding_box_vectors():
    v = load_json("reference_vectors.json")["bounding_box"]
    box = np.zeros(v["shape"])
    (r0, r1), (c0, c1) = v["box_rows"], v["box_cols"]
    box[r0:r1, c0:c1] = 1
    for case in v["cases"]:
        assert list(masking.bounding_box(box, pad=case["pad"])) == case["expect"]
        assert list(O.bounding_box(box, pad=case["pad"])) == case["expect"]


def test_reference_disk_area():
    v = load_json("reference_vectors.json")["circle_area"]
    d = synth.disk_mask(v["shape"][0], v["radius"])
    assert abs(d.sum() - np.pi * v["radius"] ** 2) <= v["rel_tol"] * d.sum()


def test_oracle_reproduces_the_small_iterate_fixtures():
    z, names = small_cases()
    assert len(names) >= 15
    for name in names:
        g = lambda k: z[f"{name}/{k}"]
        alg, sem = name.split("_")[1], name.split("_")[2]
        psi = g("guess").astype(np.complex128)
        s = g("support").astype(np.float64)
        sw = g("sw") if f"{name}/sw" in z.files else None
        for k in range(g("psi").shape[0]):
            psi, _, e = O.iteration(psi, g("modulus"), s, alg, float(g("beta")), sem)
            if sw is not None and k == int(sw[0]):
                s, _ = O.shrinkwrap_support(psi, s, float(sw[1]), float(sw[2]))
            want = g("psi")[k]
            assert np.linalg.norm(psi.astype(np.complex64) - want) <= 1e-6 * np.linalg.norm(want), \
                (name, k)
            assert abs(e - g("trace")[k]) <= 1e-12 * g("trace")[k], (name, k)
            assert np.array_equal(s.astype(np.float32), g("supports")[k]), (name, k)


def test_final_error_fixtures_cover_the_workloads():
    for wl, keys in (("cpu_disk_256", ["schedule"]), ("single_1024", ["schedule"]),
                     ("trials_512", ["raar", "dm_er_tail"])):
        d = load_json(f"final_error_{wl}.json")
        assert d["workload"] == wl and sorted(d["runs"]) == sorted(keys)
        for k in keys:
            r = d["runs"][k]
            fe = [x["final_error"] for x in r["runs"]]
            assert len(fe) == 8 and r["seeds"] == list(range(1, 9))
            assert r["final_error"]["min"] == min(fe) and r["final_error"]["max"] == max(fe)
            assert all(np.isfinite(fe)) and 0 < min(fe) < 1.5


def test_one_seed_of_the_cpu_workload_recomputed():
    d = load_json("final_error_cpu_disk_256.json")
    run = d["runs"]["schedule"]["runs"][0]
    c = synth.CONFIGS["cpu_disk_256"]
    f = synth.config_frame("cpu_disk_256", 0)
    g = O.seed_guess(256, run["seed"], f.support)
    psi, tr, _ = O.iterate(g, f.modulus, f.support, c["iterations"], c["order"], c["beta"])
    assert O.modulus_error(O.fft2(psi), f.modulus) == pytest.approx(run["final_error"], rel=1e-12)
    assert tr[-1] == pytest.approx(run["trace_last"], rel=1e-12)
