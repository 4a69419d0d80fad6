"""Summaries of rocprofv3 runs of the engine.

    python -m se_rbf_projects_amd.profiling trace <dir-with-rocpd-db>
    python -m se_rbf_projects_amd.profiling pmc <dir> [<dir> ...]
    python -m se_rbf_projects_amd.profiling traffic <fetch-dir> <write-dir> <copy-MiB> [out.json]

`trace` prints per-kernel call counts and average durations from a
`rocprofv3 --kernel-trace --stats` run; `pmc` prints per-kernel averages of the
counters of one or more `rocprofv3 --pmc ...` runs; `traffic` turns a read run and a
WRITE_SIZE run into HBM bytes per launch.  The read run is either TCC_EA0_RDREQ_128B_sum,
TCC_EA0_RDREQ_64B_sum and TCC_EA0_RDREQ_32B_sum (L2 read requests to memory by size: bytes
without a calibration; the plain 16-byte stream-copy kernel of the same run, which moves copy-MiB
in each direction, is then a check), or FETCH_SIZE, scaled by what it reads for that copy kernel
(on gfx950 FETCH_SIZE counts a 128-byte request as 64 bytes, since its TCC_BUBBLE term reads 0).
WRITE_SIZE is scaled by the copy's figure (1.0 on gfx950).  All read the rocpd SQLite database
rocprofv3 writes (the `kernels` / `counters_collection` views).
"""
import collections
import glob
import json
import os
import sqlite3
import sys


def _dbs(path):
    return sorted(glob.glob(os.path.join(path, "**", "*.db"), recursive=True))


_MANGLED_ARGS = {"f": "float", "d": "double", "i": "int", "b": "bool"}


def _demangle_simple(name):
    """kernel<args> from an Itanium-mangled serbf kernel name that the profiler left mangled
    (e.g. one with a _Float16 parameter): the nested name's last component and its simple
    template arguments (types f/d/i/b and integer literals Li<n>E)."""
    import re
    m = re.match(r"_ZN5serbf(?:12_GLOBAL__N_1)?(\d+)", name)
    if not m:
        return name
    k = m.end()
    n = int(m.group(1))
    base = name[k:k + n]
    rest = name[k + n:]
    if not rest.startswith("I"):
        return base
    args, i = [], 1
    while i < len(rest) and rest[i] != "E":
        if rest[i] == "L":
            j = rest.index("E", i)
            args.append(rest[i + 2:j])
            i = j + 1
        elif rest[i] in _MANGLED_ARGS:
            args.append(_MANGLED_ARGS[rest[i]])
            i += 1
        else:
            return base
    return base + "<" + ", ".join(args) + ">"


def _short(name):
    if name.startswith("_Z"):
        return _demangle_simple(name)
    name = name.replace("(anonymous namespace)::", "").replace("void ", "").replace("serbf::", "")
    return name.split("(")[0]


def kernel_stats(path):
    """{kernel: (calls, average_us, total_us)} from kernel-trace databases under path."""
    acc = collections.defaultdict(list)
    for db in _dbs(path):
        con = sqlite3.connect(db)
        for name, dur in con.execute("select name, duration from kernels"):
            acc[_short(name)].append(dur / 1e3)
    return {k: (len(v), sum(v) / len(v), sum(v)) for k, v in acc.items()}


def counter_stats(paths):
    """{kernel: {counter: average per dispatch}} from pmc databases."""
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    for path in paths:
        for db in _dbs(path):
            con = sqlite3.connect(db)
            for kn, cn, v in con.execute("select kernel_name, counter_name, value from counters_collection"):
                acc[_short(kn)][cn].append(v)
    return {k: {c: sum(v) / len(v) for c, v in d.items()} for k, d in acc.items()}


_RDREQ = (("TCC_EA0_RDREQ_128B", 128), ("TCC_EA0_RDREQ_64B", 64), ("TCC_EA0_RDREQ_32B", 32))


def _counter(d, name):
    for k in (name + "_sum", name):
        if k in d:
            return d[k]
    return None


def read_bytes(d):
    """Bytes one dispatch read from memory, from its L2 read requests counted by size, or None
    when the run did not count them."""
    v = [_counter(d, n) for n, _ in _RDREQ]
    if any(x is None for x in v):
        return None
    return float(sum(x * size for x, (_, size) in zip(v, _RDREQ)))


def traffic(fetch_dir, write_dir, copy_mib):
    """{kernel: {fetch_KiB, write_KiB, hbm_bytes_per_launch}} and how the read side was counted:
    read requests by size, checked on stream_copy_kernel, or FETCH_SIZE calibrated on it."""
    f = counter_stats([fetch_dir])
    w = counter_stats([write_dir])
    # the plain 16-byte copy (one vector per thread, ordinary loads and stores)
    copy = next(k for k in ("stream_copy_kernel<1, false>", "stream_copy_kernel") if k in f and k in w)
    nbytes = copy_mib * 1024 * 1024
    by_request = read_bytes(f[copy]) is not None
    if by_request:
        kf = 1.0
        fetch = {k: read_bytes(d) for k, d in f.items()}
    else:
        kf = nbytes / (f[copy]["FETCH_SIZE"] * 1024.0)
        fetch = {k: kf * d["FETCH_SIZE"] * 1024.0 for k, d in f.items() if "FETCH_SIZE" in d}
    kw = nbytes / (w[copy]["WRITE_SIZE"] * 1024.0)
    out = {}
    for k in sorted(set(f) & set(w)):
        fb, wk = fetch.get(k), w[k].get("WRITE_SIZE")
        if fb is None or wk is None:
            continue
        out[k] = {"fetch_KiB": fb / 1024.0, "write_KiB": wk, "hbm_bytes_per_launch": fb + kw * wk * 1024.0}
    # one RK4 step = stage launches 1..4 with Delta^k at the constant coefficient (STAGE 0 is the
    # right-hand-side evaluation): mvec_kernel<STAGE, HVM> (FP32 only, HVM 1), mtile2_kernel<T,
    # STAGE, HV, TPW> or stage2_kernel<T, S, NTR, STAGE, HV, U>, the first family found per precision
    for prec in ("float", "double"):
        key = ("fp32" if prec == "float" else "fp64") + " step (Delta^4 in every stage)"
        for fam, ist, ihv, pre in (("mvec_kernel", 0, 1, ""), ("mtile2_kernel", 1, 2, prec + ","),
                                   ("stage2_kernel", 3, 4, prec + ",")):
            if fam == "mvec_kernel" and prec != "float":
                continue
            st = []
            for k, val in out.items():
                if not k.startswith(fam + "<" + pre):
                    continue
                args = [x.strip() for x in k[k.index("<") + 1:k.rindex(">")].split(",")]
                if len(args) > ihv and args[ihv] == "1" and args[ist] in ("1", "2", "3", "4"):
                    st.append(val)
            if len(st) == 4 and key not in out:
                out[key] = {"hbm_bytes_per_launch": sum(x["hbm_bytes_per_launch"] for x in st) / 4,
                            "kernel": fam + "<" + (prec + ", " if pre else "") + "STAGE, ...>"}
    cal = {"copy_bytes_each_way": nbytes, "fetch_factor": kf, "write_factor": kw,
           "fetch_method": "read requests by size (128/64/32 B)" if by_request else "FETCH_SIZE x fetch_factor"}
    if by_request:
        cal["copy_read_bytes_over_moved"] = fetch[copy] / nbytes
    return {"calibration": cal, "kernels": out}


def main(argv):
    if len(argv) < 2:
        print(__doc__)
        return 1
    if argv[0] == "trace":
        st = kernel_stats(argv[1])
        tot = sum(v[2] for v in st.values())
        print("| kernel | calls | average (us) | total (us) | share |")
        print("|---|---|---|---|---|")
        for k, (c, a, t) in sorted(st.items(), key=lambda kv: -kv[1][2]):
            print(f"| `{k}` | {c} | {a:.1f} | {t:.0f} | {100 * t / tot:.1f}% |")
    elif argv[0] == "traffic":
        d = traffic(argv[1], argv[2], float(argv[3]))
        txt = json.dumps(d, indent=1)
        if len(argv) > 4:
            with open(argv[4], "w") as fh:
                fh.write(txt + "\n")
        print(txt)
    else:
        st = counter_stats(argv[1:])
        names = sorted({c for d in st.values() for c in d})
        print("| kernel | " + " | ".join(names) + " |")
        print("|---" * (len(names) + 1) + "|")
        for k, d in sorted(st.items()):
            print(f"| `{k}` | " + " | ".join(f"{d.get(c, float('nan')):.4g}" for c in names) + " |")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
