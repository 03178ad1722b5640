"""Extract the basis-set data this package ships from the reference's basis/ dir.

The reference keeps Basis Set Exchange JSON files in ``basis/*.json``
(read by ``src/desc/basis.cpp:278-360``).  The GPU host has no copy of the
reference, so the three sets used by the DF-HF / SOS-MP2 path are trimmed to
elements H..Ne and committed as data under chem_amd/basis/.

Usage: python tests/golden/extract_basis.py /root/reference/basis
"""
import json
import os
import sys

SETS = ["cc-pvdz", "cc-pvdz-jkfit", "cc-pvdz-ri"]
ELEMENTS = [str(z) for z in range(1, 11)]


def main(src):
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..", "chem_amd", "basis")
    for name in SETS:
        with open(os.path.join(src, name + ".json")) as f:
            data = json.load(f)
        out = {
            "name": data.get("name", name),
            "source": "Basis Set Exchange JSON, elements 1-10",
            "elements": {
                z: {"electron_shells": data["elements"][z]["electron_shells"]}
                for z in ELEMENTS
                if z in data["elements"]
            },
        }
        with open(os.path.join(here, name + ".json"), "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/basis")
