"""Write tests/golden/param_blocks.json: the 36 (row block, column block, tensor row, tensor
column, index set) entries of light_locked_Huang/Construct_parameter_mtx.m:13-312, read
mechanically from the text of that file (run in the build container, where the reference is
present; the JSON it writes is the fixture, the reference is not needed afterwards).

Each entry of the file is four statements:
    tmp_x                       = value_air(a,b) * ones(n,1);
    tmp_x(<index set>,1)        = anisotropic_mtx(a,b);
    i_idx(...)                  = (r*n+1:(r+1)*n)';
    j_idx(...)                  = (c*n+1:(c+1)*n)';
Index sets: face_edge_idx_in.X / .Y / .Z -> 0 / 1 / 2, vertex_Cubic_idx_in -> 3.
"""
import json
import os
import re
import sys

SRC = "/root/reference/light_locked_Huang/Construct_parameter_mtx.m"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "param_blocks.json")
SETS = {"face_edge_idx_in.X": 0, "face_edge_idx_in.Y": 1, "face_edge_idx_in.Z": 2, "vertex_Cubic_idx_in": 3}


def block_of(rng):
    """'(1:n)' -> 0, '(n+1:2*n)' -> 1, '(4*n+1:5*n)' -> 4"""
    m = re.fullmatch(r"\((?:(\d*)\*?n\+)?1:(?:(\d+)\*)?n\)'", rng.replace(" ", ""))
    if not m:
        raise ValueError(rng)
    lo = 0 if m.group(1) is None else int(m.group(1) or 1)   # '1' -> 0, 'n+1' -> 1, '4*n+1' -> 4
    hi = int(m.group(2) or 1)
    if hi != lo + 1:
        raise ValueError(rng)
    return lo


def main(src=SRC):
    entries, cur = [], {}
    for line in open(src):
        line = line.split("%")[0].strip()
        m = re.match(r"tmp_x\s*=\s*value_air\((\d),(\d)\)", line)
        if m:
            cur = {"air": (int(m.group(1)) - 1, int(m.group(2)) - 1)}
            continue
        m = re.match(r"tmp_x\(([\w.]+),1\)\s*=\s*anisotropic_mtx\((\d),(\d)\)", line)
        if m:
            cur["set"] = SETS[m.group(1)]
            cur["ab"] = (int(m.group(2)) - 1, int(m.group(3)) - 1)
            continue
        m = re.match(r"i_idx\(.*?\)\s*=\s*(.*);", line)
        if m:
            cur["row"] = block_of(m.group(1))
            continue
        m = re.match(r"j_idx\(.*?\)\s*=\s*(.*);", line)
        if m:
            cur["col"] = block_of(m.group(1))
            if cur["air"] != cur["ab"]:
                raise ValueError(f"value_air and anisotropic_mtx entries differ: {cur}")
            entries.append([cur["row"], cur["col"], cur["ab"][0], cur["ab"][1], cur["set"]])
            cur = {}
    if len(entries) != 36:
        raise ValueError(f"expected 36 entries, found {len(entries)}")
    with open(OUT, "w") as f:
        json.dump({"source": "light_locked_Huang/Construct_parameter_mtx.m:13-312",
                   "fields": ["row_block", "col_block", "tensor_row", "tensor_col", "index_set"],
                   "entries": entries}, f, indent=1)
    print(f"wrote {OUT}: {len(entries)} entries")


if __name__ == "__main__":
    main(*sys.argv[1:])
