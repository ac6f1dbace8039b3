"""Writes tests/golden/layout_kats.json: the known answers of the reference's
layout tests, as data.

Run in a container that holds the reference at REF (never on a machine
without it; the tests read only the JSON).  The extractor and node-data
tests assert many values in one repeated pattern, so those are read from
their source text with regular expressions; the results are data only
(sizes, index ranges, counts).  The handful of values in the other tests
(block size calculator, block_local_index creator, relaxed supernode, perm,
reordering) are written below by hand, each with its source line range.

    python tests/golden/transcribe_layout_kats.py
"""
import json
import os
import re

REF = os.environ.get("MFD_REFERENCE", "/root/reference")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "layout_kats.json")

EXTRACTORS = {
    "supernode_matrix": "src/test/factors/supernode_matrix_extractor_test.f90",
    "work_matrix": "src/test/factors/work_matrix_extractor_test.f90",
    "border_matrix": "src/test/factors/border_matrix_extractor_test.f90",
    "supernode_array": "src/test/right_hand/supernode_array_extractor_test.f90",
    "work_array": "src/test/right_hand/work_array_extractor_test.f90",
    "border_array": "src/test/right_hand/border_array_extractor_test.f90",
}

NODE_DATA = re.compile(r"create_node_data\(\[([\d,\s]+)\],\s*\[([\d,\s]+)\],\s*(\d+)\)")
SIZE = re.compile(r"estimate_size\((\d+)\)")
SIZE_VAL = re.compile(r'assert_equal\("estimate_size",\s*array_size,\s*(\d+)\)')
PTR = re.compile(r"get_ptr\(array,\s*([\d,\s]+)\),\s*\[\(dble\(i\),\s*i=(\d+),(\d+)\)\]")


def ints(s):
    return [int(x) for x in s.replace(" ", "").split(",") if x]


def extractor_cases(path):
    """One case per create_node_data: per-node sizes and the 1-based
    (first, last) range of each get_ptr."""
    cases, cur, pending = [], None, None
    with open(path) as f:
        for line in f:
            m = NODE_DATA.search(line)
            if m:
                cur = {"c": ints(m.group(1)), "w": ints(m.group(2)), "nb": int(m.group(3)), "sizes": {},
                       "ptrs": []}
                cases.append(cur)
                continue
            m = SIZE.search(line)
            if m:
                pending = int(m.group(1))
                continue
            m = SIZE_VAL.search(line)
            if m:
                cur["sizes"][str(pending)] = int(m.group(1))
                continue
            m = PTR.search(line)
            if m:
                cur["ptrs"].append(ints(m.group(1)) + [int(m.group(2)), int(m.group(3))])
    return cases


ND_SCALAR = re.compile(r"add_test\(\"[^\"]*\",\s*node_data%(\w+)\(([\d,\s]+)\),\s*(\d+)\)")
ND_ARRAY = re.compile(r"assert_equal\(\"(\w+)\",\s*(?:tmp_[il]|\[.*?\]),\s*\[([^\]]*)\]\)")


def node_data_cases(path):
    """node_data_test.f90: per create_node_data, the per-node arrays
    (divisible, border sizes, ...) and the scalar getters with arguments."""
    cases, cur = [], None
    with open(path) as f:
        for line in f:
            m = NODE_DATA.search(line)
            if m:
                cur = {"c": ints(m.group(1)), "w": ints(m.group(2)), "nb": int(m.group(3)), "arrays": {},
                       "calls": []}
                cases.append(cur)
                continue
            m = re.search(r"supernode_size = \[([\d,\s]+)\]", line)
            if m:
                pend_c = ints(m.group(1))
                continue
            m = re.search(r"work_size = \[([\d,\s]+)\]", line)
            if m:
                pend_w = ints(m.group(1))
                continue
            m = re.search(r"create_node_data\(supernode_size, work_size, (\d+)\)", line)
            if m:
                cur = {"c": pend_c, "w": pend_w, "nb": int(m.group(1)), "arrays": {}, "calls": []}
                cases.append(cur)
                continue
            m = ND_SCALAR.search(line)
            if m:
                cur["calls"].append([m.group(1)] + ints(m.group(2)) + [int(m.group(3))])
                continue
            m = ND_ARRAY.search(line)
            if m:
                vals = [v.strip() for v in m.group(2).split(",")]
                conv = [1 if v == ".true." else 0 if v == ".false." else int(v) for v in vals]
                cur["arrays"][m.group(1)] = conv
    return cases


HAND = {
    "block_size_calculator": {
        "source": "src/test/factors/block_size_calculator_test.f90:6-16",
        "format": "[idx, nb, n, first_block or null, expected]",
        "cases": [[1, 4, 10, None, 4], [2, 4, 10, None, 4], [3, 4, 10, None, 2],
                  [1, 4, 10, 1, 1], [2, 4, 10, 1, 4], [3, 4, 10, 1, 4], [4, 4, 10, 1, 1],
                  [1, 4, 10, 0, 4], [2, 4, 10, 0, 4], [3, 4, 10, 0, 2]],
    },
    "block_local_index_creator": {
        "source": "src/test/block_local_index_creator_test.f90:21-91",
        "cases": [
            {"local_index_lengths": [4, 0], "local_index": [2, 3, 7, 8], "c": [5, 8], "w": [4, 0], "nb": 3,
             "num_blocks": [3], "num_indices": [1, 1, 2], "rebuild_val": [2, 3, 1, 2]},
            {"local_index_lengths": [3, 2, 2, 1, 0], "local_index": [2, 4, 6, 1, 5, 2, 3, 1],
             "c": [3, 2, 4, 1, 3], "w": [3, 2, 2, 1, 0], "nb": 2, "num_blocks": [3, 2, 2, 1],
             "num_indices": [1, 1, 1, 1, 1, 1, 1, 1], "rebuild_val": [2, 2, 2, 1, 1, 2, 1, 1]},
            {"local_index_lengths": [3, 2, 3, 1, 0], "local_index": [2, 4, 7, 1, 6, 1, 2, 3, 1],
             "c": [3, 2, 4, 1, 3], "w": [3, 2, 3, 1, 0], "nb": 2, "num_blocks": [3, 2, 2, 1],
             "num_indices": [1, 1, 1, 1, 1, 2, 1, 1], "rebuild_val": [2, 2, 1, 1, 2, 1, 2, 1, 1]},
        ],
    },
    "relaxed_supernode": {
        "source": "src/test/data_structure_creators/relaxed_supernode_test.f90:16-106 "
                  "(fundamental node sizes [1,1,2,1,1,1,2] and column counts [2,2,2,2,2,2,0] from "
                  "src/test_mod/sparse_matrix_maker_m.f90:74-98)",
        "node_size": [1, 1, 2, 1, 1, 1, 2],
        "cc_fundamental": [2, 2, 2, 2, 2, 2, 0],
        "cases": [
            {"merges": [[1, 2], [4, 5], [6, 7]], "sons": [[2, 7], [3, 7], [5, 7]],
             "map": [2, 3, 5, 7], "first_node": [1, 3, 5, 7, 10], "perm": [1, 2, 3, 4, 5, 6, 7, 8, 9],
             "cc": [2, 2, 2, 0], "num_child": [0, 0, 0, 3]},
            {"merges": [[1, 2], [4, 5], [3, 6], [2, 7]], "sons": [[5, 6], [6, 7]],
             "map": [5, 6, 7], "first_node": [1, 3, 6, 10], "perm": [5, 6, 3, 4, 7, 1, 2, 8, 9],
             "cc": [2, 2, 0], "num_child": [0, 1, 1]},
            {"merges": [[1, 2], [2, 3], [3, 4], [4, 5], [5, 6], [6, 7]], "sons": [],
             "map": [7], "first_node": [1, 10], "perm": [1, 2, 3, 4, 5, 6, 7, 8, 9],
             "cc": [0], "num_child": [0]},
        ],
    },
    "perm": {
        "source": "src/test/data_structure_creators/perm_test.f90:16-27 "
                  "(perm from src/test_mod/perm_maker_m.f90:5-15)",
        "perm": [1, 7, 2, 4, 3, 5, 6, 8, 9],
        "iperm": [1, 3, 5, 4, 6, 7, 2, 8, 9],
        "permutate_of_1_to_9": [1, 7, 2, 4, 3, 5, 6, 8, 9],
        "inverse_permutate_of_1_to_9": [1, 3, 5, 4, 6, 7, 2, 8, 9],
    },
    "reordering": {
        "source": "src/test/data_structure_creators/reordering_test.f90:15-47 "
                  "(matrix from src/test_mod/sparse_matrix_maker_m.f90 make_ccs, the 9x9 fixture)",
        "ccs_perm_postorder": [1, 2, 3, 18, 19, 4, 5, 6, 7, 11, 12, 8, 9, 10, 13, 14, 15, 16, 17, 20, 21],
        "origin_val": [11, 71, 81, 22, 42, 62, 92, 33, 53, 83, 44, 94, 55, 65, 85, 66, 96, 77, 97, 88, 99],
        "ordered_val": [11, 71, 81, 77, 97, 22, 42, 62, 92, 44, 94, 33, 53, 83, 55, 65, 85, 66, 96, 88, 99],
        "small": {"num_row": [3, 2, 1], "row": [1, 2, 3, 2, 3, 3], "perm": [3, 1, 2], "iperm": [2, 3, 1],
                  "out_num_row": [3, 2, 1], "out_row": [1, 2, 3, 2, 3, 3], "ccs_perm": [6, 3, 5, 1, 2, 4]},
    },
}


def main():
    out = {"_about": "Known answers of the reference's layout and structure tests, as data (1-based). "
                     "Extractor and node-data cases are read from the test sources by "
                     "tests/golden/transcribe_layout_kats.py; the others are written there by hand."}
    out["extractors"] = {k: {"source": v, "cases": extractor_cases(os.path.join(REF, v))}
                         for k, v in EXTRACTORS.items()}
    nd = "src/test/node_data_test.f90"
    out["node_data"] = {"source": nd, "cases": node_data_cases(os.path.join(REF, nd))}
    out.update(HAND)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(OUT)


if __name__ == "__main__":
    main()
