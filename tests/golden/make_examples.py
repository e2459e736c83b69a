"""Writes tests/golden/examples/ from the reference's shipped examples.

Run in the build container only (it reads the reference tree):
    python tests/golden/make_examples.py

For each directory under the reference's examples/ it stores the
configuration's key lines (comments and blank lines dropped) as <name>.in, the
numbers of the shipped run log (examples/<name>/vmachine.log) in
expected.json (the per-iteration table and the final block), and the density
files the example ships (<name>/onebody2.dat etc.) verbatim.  The logs are the
reference's own outputs for those inputs; the engine's `vmachine` runs are
compared against them statistically (different random streams).
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "examples")
SRC = "/root/reference/examples"

DATA_FILES = {"onebody.dat", "onebody2.dat", "twobody.dat"}

FINAL = {
    "Total energy": "energy",
    "Kinetic energy": "kinetic",
    "External energy": "external",
    "Interaction energy": "interaction",
}


def parse_header(text):
    """The information blocks printed before the iteration table: a list of
    [title or key, value] in order (titles with value None)."""
    out = []
    for line in text.splitlines():
        if line.startswith("Step  Energy"):
            break
        s = line.strip()
        if not s or set(s) <= set("=_\\/|. `()-"):
            continue
        if s.isupper():
            out.append([s, None])
        elif ":" in s:
            k, v = s.split(":", 1)
            out.append([k.strip(), v.strip()])
    return out


def parse_log(text):
    rows, final = [], {}
    for line in text.splitlines():
        parts = line.split()
        if len(parts) == 11 and re.fullmatch(r"\d+", parts[0]):
            try:
                rows.append([int(parts[0])] + [float(p) for p in parts[1:]])
            except ValueError:
                pass
            continue
        m = re.match(r"\s*(Total|Kinetic|External|Interaction) energy\s*:\s*(\S+) \(with STD = (\S+)\)",
                     line)
        if m:
            key = FINAL[m.group(1) + " energy"]
            final[key] = float(m.group(2))
            final[key + "_std"] = float(m.group(3))
        m = re.match(r"\s*Acceptence Ratio\s*:\s*(\S+)", line)
        if m:
            final["acceptance"] = float(m.group(1))
        m = re.match(r"Average CPU time before convergence:\s*(\S+)", line)
        if m:
            final["cpu_time_per_iteration"] = float(m.group(1))
        m = re.match(r"Elapsed time:\s*(\S+)s", line)
        if m:
            final["elapsed_s"] = float(m.group(1))
        m = re.match(r"Running on (\d+) CPU threads", line)
        if m:
            final["mpi_ranks"] = int(m.group(1))
    return rows, final


def main():
    if not os.path.isdir(SRC):
        print("reference examples not found at", SRC)
        return 1
    os.makedirs(OUT, exist_ok=True)
    expected = {}
    for name in sorted(os.listdir(SRC)):
        d = os.path.join(SRC, name)
        cfg = os.path.join(d, "input.in")
        log = os.path.join(d, "vmachine.log")
        if not (os.path.isfile(cfg) and os.path.isfile(log)):
            continue
        with open(cfg) as f:
            keys = [ln.rstrip() for ln in f if ln.strip() and not ln.lstrip().startswith("#")]
        with open(os.path.join(OUT, name + ".in"), "w") as f:
            f.write("\n".join(keys) + "\n")
        with open(log, encoding="utf-8", errors="replace") as f:
            text = f.read()
        rows, final = parse_log(text)
        expected[name] = {"rows": rows, "final": final, "header": parse_header(text)}
        # output data files the example ships beside its log (e.g. the spatial
        # density onebody2.dat), verbatim
        for fn in sorted(os.listdir(d)):
            if fn.endswith(".dat") and fn in DATA_FILES:
                os.makedirs(os.path.join(OUT, name), exist_ok=True)
                with open(os.path.join(d, fn), "rb") as src, \
                        open(os.path.join(OUT, name, fn), "wb") as dst:
                    dst.write(src.read())
    with open(os.path.join(OUT, "expected.json"), "w") as f:
        json.dump(expected, f, indent=1)
    print("wrote", len(expected), "examples to", OUT)
    return 0


if __name__ == "__main__":
    sys.exit(main())
