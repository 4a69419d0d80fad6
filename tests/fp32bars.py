"""FP32 error bars of the GPU parity tests.

Each bar sits about ten times above the relative error measured on an MI355X for
the same case (the measured value is written beside the bar), so a regression of
an order of magnitude fails.  With SERBF_F32_LOG=<file> set, every check appends
its measured error to that file as a JSON line."""
import json
import os


def check(name, err, bar):
    log = os.environ.get("SERBF_F32_LOG")
    if log:
        with open(log, "a") as f:
            f.write(json.dumps({"case": name, "err": float(err), "bar": bar}) + "\n")
    assert err < bar, f"{name}: FP32 relative error {err:.3e} above its bar {bar:.1e}"
