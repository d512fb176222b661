#!/usr/bin/env python3
"""Write tests/golden/zig_reference_tables.json from the reference's table text.

The reference embeds its ziggurat tables as 19-significant-digit decimal
literals (src/dist/ziggurat_tables.rs:6 for R, :8-266 for X[257], :268-526 for
Y[257]).  Each literal rounds to a unique binary64; this script parses them
(Python float() rounds correctly, like Rust's f64 parser) and stores the
values as hex floats, so the committed fixture is plain data that the tests
compare the generated tables against without needing the reference.

Run in a container where the reference is mounted:
    python tests/golden/make_zig_fixture.py /root/reference
"""
import json
import os
import re
import sys


def parse(path):
    text = open(path).read()
    num = r"([-+]?\d\.\d+e[-+]\d+)"
    r = float(re.search(r"ZIGGURAT_R: f64 = " + num, text).group(1))

    def block(name):
        body = text.split(f"pub const {name}")[1].split("];")[0]
        return [float(v) for v in re.findall(num + r",\s*// ", body)]

    X, Y = block("ZIGGURAT_X"), block("ZIGGURAT_Y")
    assert len(X) == 257 and len(Y) == 257, (len(X), len(Y))
    return r, X, Y


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    r, X, Y = parse(os.path.join(ref, "src/dist/ziggurat_tables.rs"))
    out = {
        "source": "src/dist/ziggurat_tables.rs (R :6, X :8-266, Y :268-526)",
        "R": r.hex(),
        "X": [x.hex() for x in X],
        "Y": [y.hex() for y in Y],
    }
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "zig_reference_tables.json")
    with open(dst, "w") as fh:
        json.dump(out, fh, indent=0)
    print("wrote", dst)


if __name__ == "__main__":
    main()
