"""Writes tests/golden/donnelly_kats.json: the known-answer vectors of the reference's
Donnelly stem layout tests, read from its test source (data only: each case's input path
and expected stem index, with the file and line it comes from).

    python tests/golden/extract_donnelly_kats.py [/root/reference]

Cases (src/stem_strategy/donnelly/core.rs, block height 3):
  * donnelly_core_get_child_idx_produces_correct_values: a path of left/right choices
    from the root -> the stem index reached (traverse);
  * donnelly_core_get_both_child_idxs_produces_correct_values: a path -> the stem indices
    of the left and right children of the node reached;
  * donnelly_core_get_child_idx_unrolled_produces_correct_values: the same paths through
    traverse_head / traverse_tail;
  * donnelly_core_traverse_block_produces_correct_values: a sequence of 3-level block
    choices (child 0..7) -> the stem index reached.
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
SRC = "src/stem_strategy/donnelly/core.rs"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "donnelly_kats.json")


def main():
    with open(os.path.join(REF, SRC)) as f:
        lines = f.read().split("\n")
    groups, cur = {}, []
    case = re.compile(r"#\[case\(vec!\[(.*?)\],\s*(\(?[\d, ]+\)?)\)\]")
    for no, line in enumerate(lines, 1):
        m = case.search(line)
        if m:
            items = [x.strip() for x in m.group(1).split(",") if x.strip()]
            want = [int(x) for x in re.findall(r"\d+", m.group(2))]
            cur.append({"line": no, "input": items, "expect": want})
            continue
        fn = re.search(r"fn (donnelly_core_\w+)\(", line)
        if fn and cur:
            groups[fn.group(1)] = cur
            cur = []
    out = {"source": SRC, "block_height": 3, "groups": {}}
    for name, cases in groups.items():
        conv = []
        for c in cases:
            if c["input"] and c["input"][0] in ("true", "false"):
                inp = [x == "true" for x in c["input"]]
            else:
                inp = [int(x) for x in c["input"]]
            conv.append({"line": c["line"], "input": inp,
                         "expect": c["expect"] if len(c["expect"]) > 1 else c["expect"][0]})
        out["groups"][name] = conv
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print({k: len(v) for k, v in out["groups"].items()})


if __name__ == "__main__":
    main()
