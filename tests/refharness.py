"""Runs the reference solver's own driver-level sources on given inputs.

oracle/_ref/ref_harness is built by `make -C oracle ref` from the unmodified
reference files wrappers.c, test_cases.c, layers.c, patches.c and debug.c
(Advection/C_MPI_OMP_ACC/code) around oracle/ref_harness.c.  It exists only where
the reference tree was present at build time; callers skip when it is absent.

Test infrastructure: imported by tests/ only."""
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")

INT_ARRAYS = {"p.pid2gid", "p.gid2pid", "p.part_pids", "p.nodeset.idx"}


def available():
    return os.path.exists(HARNESS)


def adv_env(nodeset_file="NONE", num_layers=8, stencil_size=31, model_height=12000.0, timestep=1800.0,
            test_case=2, from_file=1, use_metis=1):
    """The ADV_* environment get_rt_config reads (wrappers.c:346-395)."""
    return {"ADV_NODESET_FROM_FILE": str(from_file), "ADV_NODESET_FILE": nodeset_file,
            "ADV_USE_METIS": str(use_metis), "ADV_TEST_CASE": str(test_case),
            "ADV_STENCIL_SIZE": str(stencil_size), "ADV_NUM_LAYERS": str(num_layers),
            "ADV_MODEL_HEIGHT": repr(float(model_height)), "ADV_TIMESTEP_LENGTH": repr(float(timestep))}


class Result:
    def __init__(self, out_dir, stdout):
        self.dir = out_dir
        self.stdout = stdout
        self.scalars = {}
        with open(os.path.join(out_dir, "scalars.txt")) as f:
            for line in f:
                k, _, v = line.rstrip("\n").partition(" ")
                self.scalars[k] = v

    def s(self, name, kind=float):
        return kind(self.scalars[name])

    def a(self, name):
        dt = np.int32 if name in INT_ARRAYS else np.float64
        return np.fromfile(os.path.join(self.dir, name), dtype=dt)


def run(xyz5, idx, D, env, times=(), samples=None, qt=None):
    """Run the reference sequence get_rt_config -> init_phys_constants -> (node set and
    stencils given) -> init_global_layers -> (one-part partition) -> init_local_patch ->
    get_model_ICs, then update_U at `times`, the TC2 point functions at `samples`
    ((h, phi) rows) and verify_solution on the state `qt` (N x Nv)."""
    work = tempfile.mkdtemp(prefix="refh_")
    ind, outd = os.path.join(work, "in"), os.path.join(work, "out")
    os.makedirs(ind)
    os.makedirs(outd)
    xyz5 = np.asarray(xyz5, dtype=np.float64)
    np.ascontiguousarray(xyz5.T).tofile(os.path.join(ind, "nodes.bin"))
    np.ascontiguousarray(idx, dtype=np.int32).tofile(os.path.join(ind, "idx.bin"))
    np.ascontiguousarray(D, dtype=np.float64).tofile(os.path.join(ind, "D.bin"))
    if len(times):
        np.asarray(times, dtype=np.float64).tofile(os.path.join(ind, "times.bin"))
    if samples is not None:
        np.ascontiguousarray(samples, dtype=np.float64).tofile(os.path.join(ind, "samples.bin"))
    if qt is not None:
        np.ascontiguousarray(qt, dtype=np.float64).tofile(os.path.join(ind, "qt.bin"))
    full_env = {k: v for k, v in os.environ.items() if not k.startswith("ADV_")}
    full_env.update(env)
    r = subprocess.run([HARNESS, ind, outd], env=full_env, capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        raise RuntimeError(f"ref_harness failed ({r.returncode}):\n{r.stdout}\n{r.stderr}")
    return Result(outd, r.stdout)


def parse_verification(text):
    """{min, max, L1, L2, Linf, mass_real, mass_approx} from verify_solution's banner
    (wrappers.c:165-171)."""
    import re

    pats = {"min": r"Min\(q\) = (\S+)", "max": r"Max\(q\) = (\S+)", "L1": r"L1\(q\) = (\S+)",
            "L2": r"L2\(q\) = (\S+)", "Linf": r"Linf\(q\) = (\S+)",
            "mass_real": r"Total Mass \(real\) -> (\S+)", "mass_approx": r"Total Mass \(approximated\) -> (\S+)"}
    return {k: float(re.search(p, text).group(1)) for k, p in pats.items()}
