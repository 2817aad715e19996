"""Extract the reference's own known answers for this path into a fixture.

The reference holds no unit tests or golden vectors for the OSTR path; the only
known answers are printed outputs of its notebook driver (ostr.ipynb, code
cell 3, raw JSON lines 1271-1275): the OS regroup order for 25 views in 5
subsets and the array sizes of a 25-view 3584x1400 scan. This script reads the
notebook as JSON text (nothing is executed) and writes reference_outputs.json.
Run it in a container that has the reference mounted:
    python tests/golden/make_golden.py /root/reference/ostr.ipynb
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(nb_path):
    nb = json.load(open(nb_path))
    texts = []
    for cell in nb["cells"]:
        for out in cell.get("outputs", []):
            t = out.get("text", "")
            texts.append("".join(t) if isinstance(t, list) else t)
    text = "\n".join(texts)
    m = re.search(r"Indexes are\s+\[([\d\s]+)\]", text)
    order = [int(v) for v in m.group(1).split()]
    sizes = sorted({int(v) for v in re.findall(r"\((\d+),\)", text)})
    sub = int(re.search(r"sub_b_size is\s+(\d+)", text).group(1))
    fixture = {
        "source": "ostr.ipynb printed outputs (raw lines 108, 1271-1275)",
        "n_views": 25, "n_subsets": 5, "bins_x": 3584, "bins_y": 1400,
        "regroup_order": order,
        "all_b_size": sizes,
        "sub_b_size_printed": sub,
    }
    with open(os.path.join(HERE, "reference_outputs.json"), "w") as f:
        json.dump(fixture, f, indent=1)
    print(fixture)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/ostr.ipynb")
