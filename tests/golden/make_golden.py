"""Extract the golden fixtures of the RBF-FD advection tests from the reference's
own data files.

Run once in a container that holds the reference (its path is the first
argument, default /root/reference) with an interpreter that has h5py:

    /opt/conda/bin/python3.9 tests/golden/make_golden.py /root/reference

Writes into tests/golden/:
  icos_<N>.npy        float64 [N, 5] (x, y, z, lambda, phi) of nodesets/ICOS_<N>.nc
                      for N = 42, 162, 642, 2562, 10242 (each level's first
                      nodes repeat the previous level to 1e-15)
  icos_40962_moments.json  order-independent moments of nodesets/ICOS_40962.nc
  md25.npy            float64 [25, 4] columns x y z w of nodesets/md.00025
  md25_knn10.npy      int32 [25, 10]: the stencil table the reference prototype
                      printed for MD_25 with n = 10 (Advection/Python/Solver.ipynb:228-252)
  ICOS_42.nc, ICOS_642.nc, MD_25.nc
                      three of the reference's node files, copied unchanged, as inputs of
                      the node-file reader's tests
  md25_nc.npy         float64 [25, 5] (x, y, z, lambda, phi) of nodesets/MD_25.nc

Only data is read (HDF5 datasets, a text table, and one notebook output cell
parsed as JSON); no reference code is imported or executed.
"""
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def read_nc(path):
    import h5py

    with h5py.File(path, "r") as f:
        g = f["nodeset"]
        return np.stack([np.asarray(g[k][:], dtype=np.float64) for k in ("x", "y", "z", "lambda", "phi")], axis=1)


def moments(xyz):
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    out = {"N": int(len(x))}
    for a in range(5):
        for b in range(5 - a):
            for c in range(5 - a - b):
                if a + b + c == 0:
                    continue
                out[f"m{a}{b}{c}"] = float(np.sum(x**a * y**b * z**c))
    out["abs_sum"] = float(np.sum(np.abs(x) + 2.0 * np.abs(y) + 3.0 * np.abs(z)))
    return out


def knn_table_from_notebook(path):
    nb = json.load(open(path))
    for cell in nb["cells"]:
        src = "".join(cell.get("source", []))
        if cell.get("cell_type") != "code" or src.strip().split("\n")[-1].strip() != "idx":
            continue
        for o in cell.get("outputs", []):
            txt = "".join(o.get("data", {}).get("text/plain", []))
            if txt.startswith("array([[ 0, 14"):
                body = txt[len("array(") : txt.rindex(")")]
                rows = json.loads(body.replace("\n", "").replace(" ", ""))
                return np.asarray(rows, dtype=np.int32)
    raise RuntimeError("stencil table output not found")


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    ns = os.path.join(ref, "nodesets")
    prev = None
    for N in (42, 162, 642, 2562, 10242):
        a = read_nc(os.path.join(ns, f"ICOS_{N}.nc"))
        if prev is not None:
            assert np.abs(a[: len(prev)] - prev).max() <= 1e-15, f"ICOS_{N} does not extend the previous level"
        np.save(os.path.join(HERE, f"icos_{N}.npy"), a)
        prev = a
    big = read_nc(os.path.join(ns, "ICOS_40962.nc"))
    assert np.abs(big[: len(prev)] - prev).max() <= 1e-15
    with open(os.path.join(HERE, "icos_40962_moments.json"), "w") as f:
        json.dump(moments(big[:, :3]), f, indent=1)
    md = np.loadtxt(os.path.join(ns, "md.00025"))
    np.save(os.path.join(HERE, "md25.npy"), md)
    tab = knn_table_from_notebook(os.path.join(ref, "Advection", "Python", "Solver.ipynb"))
    assert tab.shape == (25, 10)
    np.save(os.path.join(HERE, "md25_knn10.npy"), tab)
    for name in ("ICOS_42.nc", "ICOS_642.nc", "MD_25.nc"):
        shutil.copyfile(os.path.join(ns, name), os.path.join(HERE, name))
    np.save(os.path.join(HERE, "md25_nc.npy"), read_nc(os.path.join(ns, "MD_25.nc")))
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
