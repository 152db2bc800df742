"""Writes tests/golden/functors_ref.json: inputs and outputs of CTLAB's own
SART_Weighting<float>::Proj/Back (src/SARTWeighting.h) and FISTA
(src/FastIterativeShrinkageAlgorithm.h), produced by oracle/_ref/ref_probe,
which `make -C oracle ref` compiles from the headers in the reference tree."""
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

if __name__ == "__main__":
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = subprocess.run([os.path.join(ROOT, "oracle", "_ref", "ref_probe")], check=True,
                         capture_output=True, text=True).stdout
    data = json.loads(out)
    data["source"] = "oracle/_ref/ref_probe (CTLAB src/SARTWeighting.h, src/FastIterativeShrinkageAlgorithm.h)"
    with open(os.path.join(HERE, "functors_ref.json"), "w") as f:
        json.dump(data, f, indent=0)
    print("wrote functors_ref.json")
