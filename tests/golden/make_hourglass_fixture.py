"""Write tests/golden/hourglass_quads.json: the corner points of the 22
plane walls that the reference program builds itself for GeometrySource = 0
(G_MakeGeometry, src/DEM/Prtcl_Geometry.f90:557-615), in the order it adds
them.  Reads the reference's source text (in this container only); the
fixture is the data, the script is how it was made:

    python3 tests/golden/make_hourglass_fixture.py /root/reference
"""
import json
import os
import re
import sys


def main(ref: str) -> None:
    text = open(os.path.join(ref, "src", "DEM", "Prtcl_Geometry.f90")).read()
    block = text[text.index("if(DEM_Opt%GeometrySource ==0)"):text.index("elseif(DEM_Opt%GeometrySource ==1)")]
    num = r"\s*([-+0-9.]+)_RK\s*"
    pts = {m.group(1): [float(m.group(k)) for k in (2, 3, 4)]
           for m in re.finditer(r"(p\d\d)\s*=\s*real3\(" + num + "," + num + "," + num + r"\)", block)}
    quads = []
    for m in re.finditer(r"add_PlaneWall\(\s*(p\d\d)\s*,\s*(p\d\d)\s*,\s*(p\d\d)\s*,\s*(p\d\d)\s*,\s*(\d+)\s*,\s*(\d+)\s*,"
                         r"\s*infinite\s*=\s*\.(\w+)\.\s*\)", block):
        quads.append(dict(corners=[pts[m.group(k)] for k in (1, 2, 3, 4)], user_id=int(m.group(5)),
                          wall_type=int(m.group(6)), infinite=m.group(7) == "true", both_side=False))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hourglass_quads.json")
    with open(out, "w") as f:
        json.dump(dict(source="src/DEM/Prtcl_Geometry.f90:557-615 (GeometrySource = 0)", quads=quads), f, indent=1)
    print(f"{len(quads)} quads from {len(pts)} points -> {out}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
