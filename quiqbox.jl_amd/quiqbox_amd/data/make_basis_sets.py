"""Extract the Gaussian-format basis-set texts shipped with Quiqbox.jl
(lib/BasisSets/BasisSets.jl, collected there from basissetexchange.org) into
basis_sets.json: {basis name: [text or null for Z = 1, 2, ...]}.

Run once in a checkout that has the reference tree; the JSON is the committed
data file that genGaussTypeOrbSeq reads.
"""
import json
import os
import re
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/lib/BasisSets/BasisSets.jl"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis_sets.json")


def main():
    text = open(SRC).read()
    names_block = text[text.index("AtomicGTOrbSetNames"):text.index("AtomicGTOrbSetTexts")]
    names = re.findall(r'"([^"]+)"', names_block)
    body = text[text.index("AtomicGTOrbSetTexts"):]
    body = body[body.index("([") + 2:]
    # split into one block per basis set: each starts with Memory{NothingOr{String}}([
    blocks = body.split("Memory{NothingOr{String}}([")[1:]
    assert len(blocks) == len(names), (len(blocks), names)
    out = {}
    for name, blk in zip(names, blocks):
        entries = []
        pos = 0
        tok = re.compile(r'"""(.*?)"""|\bnothing\b', re.S)
        for m in tok.finditer(blk):
            entries.append(m.group(1) if m.group(0).startswith('"""') else None)
        out[name] = entries
    with open(OUT, "w") as f:
        json.dump(out, f, indent=0)
    print({k: len(v) for k, v in out.items()})


if __name__ == "__main__":
    main()
