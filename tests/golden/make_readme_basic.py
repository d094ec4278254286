"""Writes tests/golden/readme_path_basic.txt: the oracle's TileBuilder calls
for the README path (src/lib.rs:35-42, examples/basic.rs:24-32: Move(400,300),
Quadratic(500,200 -> 400,100), Cubic(350,150; 100,250 -> 400,300), Close,
identity transform) in the text format examples/basic.rs:6-19 prints.

    python tests/golden/make_readme_basic.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

OUT = os.path.join(HERE, "readme_path_basic.txt")


def main():
    import oracle_ref
    from ochre_amd import workloads as W
    from parity import basic_rs_text
    v, p, w, xf = W.readme_path().path(0)
    r = oracle_ref.Rasterizer()
    r.fill(v, p, w, xf)
    text = basic_rs_text(r.finish())
    with open(OUT, "w") as f:
        f.write(text)
    print(f"{OUT}: {text.count('tile at')} tiles, {text.count('span at')} spans")


if __name__ == "__main__":
    main()
