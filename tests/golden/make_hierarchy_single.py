"""Writes tests/golden/hierarchy_single.json: the single-linkage known answers
of the reference's own tests (scipy_numba/scipy_numba/cluster/tests/
hierarchy_test_data.py: ytdist with linkage_ytdist_single, Q_X with
fcluster_maxclust), parsed from that file as text (ast.literal_eval of the
array literals; nothing in the reference is imported or run)."""
import ast
import json
import re
from pathlib import Path

SRC = Path("/root/reference/scipy_numba/scipy_numba/cluster/tests/hierarchy_test_data.py")


def _literal(text, name):
    m = re.search(rf"^{name} = array\((.*?)\)\n", text, re.S | re.M)
    return ast.literal_eval(m.group(1))


def _dict_of_arrays(text, name):
    m = re.search(rf"^{name} = \{{(.*?)\}}\n", text, re.S | re.M)
    out = {}
    for key, body in re.findall(r"([0-9.]+): array\((\[.*?\])\)", m.group(1), re.S):
        out[key] = ast.literal_eval(body)
    return out


def main():
    text = SRC.read_text()
    data = {
        "source": "scipy_numba/scipy_numba/cluster/tests/hierarchy_test_data.py (used by "
                  "test_hierarchy.py:76-84 and :172-185)",
        "ytdist": _literal(text, "ytdist"),
        "linkage_ytdist_single": _literal(text, "linkage_ytdist_single"),
        "Q_X": _literal(text, "Q_X"),
        "fcluster_maxclust": _dict_of_arrays(text, "fcluster_maxclust"),
    }
    out = Path(__file__).with_name("hierarchy_single.json")
    out.write_text(json.dumps(data, indent=1))
    print(out)


if __name__ == "__main__":
    main()
