"""Writes tests/golden/ellipse_ref.json: CTLAB's own SheppLoganPhantom table and
Ellipsoid::intersectionLength on seeded lines (src/Ellipsoid.h:209-317),
produced by oracle/_ref/ellipse_probe, which `make -C oracle ref` compiles from
the header in the reference tree."""
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

if __name__ == "__main__":
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = subprocess.run([os.path.join(ROOT, "oracle", "_ref", "ellipse_probe")], check=True,
                         capture_output=True, text=True).stdout
    data = json.loads(out)
    data["source"] = "oracle/_ref/ellipse_probe (CTLAB src/Ellipsoid.h)"
    data["columns"] = {"shepp_logan": ["cx", "cy", "a", "b", "rot", "intensity"],
                       "chords": ["cx", "cy", "a", "b", "rot", "sx", "sy", "ex", "ey", "length"]}
    with open(os.path.join(HERE, "ellipse_ref.json"), "w") as f:
        json.dump(data, f, indent=0)
    print("wrote ellipse_ref.json")
