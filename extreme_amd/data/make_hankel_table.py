"""Write the 431-tap J0/J1 digital Hankel filter (ndec = 40) as a text table.

The table is numerical data, not code: the filter weights that extrEMe ships
in Extreme.Cartesian/Resources/hankelcoeff.xml (the <Coefficients ndec="40">
block, lines 1302-1734) and loads with HankelCoefficients.LoadN40()
(Extreme.Cartesian/Green/Scalar/Impl/HankelCoefficients.cs:37,40-65).
The Green-tensor assembly needs the weights verbatim, so this script copies
the two weight columns (h0 for J0, h1 for J1) in their original order into
hankel_ndec40.txt, one "h0 h1" pair per line, each value printed with
Python's shortest round-trip repr so that strtod() recovers the same double.

Run once, in a checkout that has the reference mounted:
    python extreme_amd/data/make_hankel_table.py /root/reference
"""
import os
import sys
import xml.etree.ElementTree as ET


def main(ref_root):
    xml_path = os.path.join(ref_root, "Extreme.Cartesian", "Resources", "hankelcoeff.xml")
    root = ET.parse(xml_path).getroot()
    block = [c for c in root.findall("Coefficients") if c.get("ndec") == "40"]
    if len(block) != 1:
        raise SystemExit("expected exactly one ndec=40 block")
    lines = block[0].findall("line")
    h0 = [float(l.get("h0")) for l in lines]
    h1 = [float(l.get("h1")) for l in lines]
    if len(h0) != 431:
        raise SystemExit(f"expected 431 taps, got {len(h0)}")
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hankel_ndec40.txt")
    with open(out, "w") as f:
        f.write("# J0 (h0) and J1 (h1) weights of the 431-tap ndec=40 Hankel filter\n")
        f.write("# source: Extreme.Cartesian/Resources/hankelcoeff.xml, <Coefficients ndec=\"40\">\n")
        for a, b in zip(h0, h1):
            f.write(f"{a!r} {b!r}\n")
    print(out, len(h0))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
