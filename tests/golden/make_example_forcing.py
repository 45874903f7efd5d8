"""Writes tests/golden/example/example.meteo: the reference example watershed's station
records (benchmark/input/example/example.meteo) from 2008-12-31 00:00 to 2009-01-15 00:00,
the first two weeks of that project's simulation (START 2009-01-01 00:00 in example.para),
with the file's three header lines; and copies example.lai (the LAI forcing) and
example.bc (the boundary-condition series) whole.

Run in a container that holds the reference:  python tests/golden/make_example_forcing.py
"""
import os
import shutil

SRC = "/root/reference/benchmark/input/example"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "example")


def main():
    with open(os.path.join(SRC, "example.meteo")) as f:
        lines = f.readlines()
    head, body = lines[:3], lines[3:]
    keep = [ln for ln in body if "2008-12-31 00:00" <= ln[:16] <= "2009-01-15 00:00"]
    with open(os.path.join(DST, "example.meteo"), "w") as f:
        f.writelines(head + keep)
    for name in ("example.lai", "example.bc"):
        shutil.copyfile(os.path.join(SRC, name), os.path.join(DST, name))
    print(f"{len(keep)} meteo records, example.lai and example.bc copied")


if __name__ == "__main__":
    main()
