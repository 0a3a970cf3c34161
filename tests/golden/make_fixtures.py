"""Convert the reference's golden transcripts into JSON fixtures.

Reads tests/tp_01.output, tests/tp_02.output, tests/transfer_01.output and
tests/transfer_02.output of nlsmrg/dealii-stfem (data files of the
reference's own tests) and writes machine-readable fixtures next to this
script.  Run once in a container that has the reference checked out:

    python tests/golden/make_fixtures.py /root/reference
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def parse_matrix_transcript(text):
    """Split a print_formatted transcript into (headers, matrix) items.

    print_formatted (tp_02.cc:12-30) writes each entry as %7.2f or as seven
    blanks when |x| < 0.01, one row per line, and ends a matrix with an
    empty line.  Blank entries are stored as None.
    """
    items, headers, rows = [], [], []

    def is_row(line):
        if not line or len(line) % 7:
            return False
        for k in range(0, len(line), 7):
            c = line[k:k + 7]
            if c.strip() and not re.fullmatch(r"\s*-?\d+\.\d\d", c):
                return False
        return True

    for line in text.split("\n"):
        if is_row(line):
            rows.append([None if not line[k:k + 7].strip() else float(line[k:k + 7])
                         for k in range(0, len(line), 7)])
        elif line == "":
            if rows:
                items.append({"headers": headers, "matrix": rows})
                headers, rows = [], []
        else:
            if rows:
                items.append({"headers": headers, "matrix": rows})
                headers, rows = [], []
            headers.append(line.strip())
    if rows:
        items.append({"headers": headers, "matrix": rows})
    return items


def parse_transfer_01(text):
    """transfer_01.output: per run DoFs, levels, average iterations; tables."""
    runs = []
    cur = {}
    for line in text.split("\n"):
        m = re.match(r":: Number of active cells: (\d+)", line)
        if m:
            cur = {"cells": int(m.group(1))}
            continue
        m = re.match(r":: Number of degrees of freedom: (\d+)", line)
        if m:
            cur["dofs"] = int(m.group(1))
            continue
        m = re.match(r":: Min Level (\d+)\s+Max Level (\d+)", line)
        if m:
            cur["max_level"] = int(m.group(2))
            continue
        m = re.match(r"Average GMRES iterations: ([\d.]+) \(over (\d+) timesteps\)", line)
        if m:
            cur["avg_its"] = float(m.group(1))
            cur["slabs"] = int(m.group(2))
            runs.append(cur)
            cur = {}
    tables = []
    for block in re.findall(r"cells s-dofs.*?\n((?:\s+\d+.*\n)+)", text):
        rows = []
        for line in block.strip("\n").split("\n"):
            tok = line.split()
            nums = [t for t in tok if t != "-"]
            rows.append({"cells": int(nums[0]), "s_dofs": int(nums[1]), "t_dofs": int(nums[2]),
                         "st_dofs": int(nums[3]), "Linf": float(nums[4]),
                         "L2": float(nums[6] if len(nums) > 7 else nums[5]),
                         "H1": float(nums[8] if len(nums) > 7 else nums[6])})
        tables.append(rows)
    return {"runs": runs, "tables": tables}


def parse_tp_01(text):
    runs = []
    cur = {}
    for line in text.split("\n"):
        m = re.match(r":: Number of active cells: (\d+)", line)
        if m:
            cur = {"cells": int(m.group(1))}
            continue
        m = re.match(r":: Number of degrees of freedom: (\d+)", line)
        if m:
            cur["dofs"] = int(m.group(1))
            continue
        m = re.match(r":: Min Level (\d+)\s+Max Level (\d+)", line)
        if m:
            cur["max_level"] = int(m.group(2))
            continue
        m = re.match(r"Average GMRES iterations ([\d.]+) \((\d+) gmres_iterations / (\d+) timesteps\)",
                     line)
        if m:
            cur["avg_its"] = float(m.group(1))
            cur["total_its"] = int(m.group(2))
            cur["slabs"] = int(m.group(3))
            runs.append(cur)
            cur = {}
    tables = []
    for block in re.findall(r"Convergence table k=(\d+)\n.*?\n((?:\s+\d+.*\n)+)", text):
        k, body = block
        rows = []
        for line in body.strip("\n").split("\n"):
            nums = [t for t in line.split() if t != "-"]
            rows.append({"cells": int(nums[0]), "s_dofs": int(nums[1]), "t_dofs": int(nums[2]),
                         "st_dofs": int(nums[3]), "work": int(nums[4]),
                         "Linf": float(nums[5]),
                         "L2": float(nums[7] if len(nums) > 8 else nums[6]),
                         "H1": float(nums[9] if len(nums) > 8 else nums[7])})
        tables.append({"k": int(k), "rows": rows})
    return {"runs": runs, "tables": tables}


def main(ref):
    t = os.path.join(ref, "tests")
    out = {}
    with open(os.path.join(t, "tp_02.output")) as f:
        out["tp_02"] = parse_matrix_transcript(f.read())
    with open(os.path.join(t, "transfer_02.output")) as f:
        out["transfer_02"] = parse_matrix_transcript(f.read())
    with open(os.path.join(t, "transfer_01.output")) as f:
        out["transfer_01"] = parse_transfer_01(f.read())
    with open(os.path.join(t, "tp_01.output")) as f:
        out["tp_01"] = parse_tp_01(f.read())
    for k, v in out.items():
        path = os.path.join(HERE, k + ".json")
        with open(path, "w") as f:
            json.dump({"source": f"tests/{k}.output of nlsmrg/dealii-stfem", "data": v}, f,
                      indent=None, separators=(",", ":"))
        print("wrote", path)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
