"""Writes qop_pc_names.json: the names of the reference's per-precision C
prototypes (include/qop_pc.h) whose arguments and result carry no QDP type,
i.e. the functions a C-ABI replacement can export.  Run in a container that
holds the reference source at /root/reference; the JSON is the committed
fixture (names only) that tests/test_abi.py checks the library against."""
import json
import os
import re

SRC = "/root/reference/include/qop_pc.h"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "qop_pc_names.json")


def main():
    text = open(SRC).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"^#(?:[^\n]*\\\n)*[^\n]*", "", text, flags=re.M)  # macros, continued lines too
    names = []
    for decl in text.split(";"):
        m = re.search(r"\b(QOP_PC_\w+)\s*\(([^)]*)\)\s*$", decl.strip(), flags=re.S)
        if not m:
            continue
        if "QDP_" in m.group(2) or "QDP_" in decl[: m.start(1)]:
            continue
        names.append(m.group(1)[len("QOP_PC_"):])
    json.dump({"source": "include/qop_pc.h", "non_qdp": names}, open(OUT, "w"), indent=1)
    print(len(names), "names ->", OUT)


if __name__ == "__main__":
    main()
