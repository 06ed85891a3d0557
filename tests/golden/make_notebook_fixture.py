"""Extract the reference notebook's stored outputs into a JSON fixture.

Reads jax/taylor_green_colab.ipynb (a data file of the reference: its stored
cell outputs are the only published numbers for this path) and writes
tests/golden/notebook_tgv_n128.json.  Run in the development container:

    python tests/golden/make_notebook_fixture.py /root/reference
"""
import json
import os
import re
import sys


def main(ref_root):
    nb = json.load(open(os.path.join(ref_root, "jax", "taylor_green_colab.ipynb")))
    text = ""
    for cell in nb["cells"]:
        for out in cell.get("outputs", []):
            if "text" in out:
                text += "".join(out["text"])
    ens = {}
    for m in re.finditer(r"time=\s*([0-9.]+)\s+enstrophy=([0-9.eE+-]+)", text):
        ens[float(m.group(1))] = m.group(2)
    final_energy = re.search(r"Final energy: ([0-9.eE+-]+)", text).group(1)
    max_div = re.search(r"Maximum divergence: ([0-9.eE+-]+)", text).group(1)
    run = re.search(r"Simulated time: ([0-9.]+) \((\d+) steps\) in ([0-9.]+) s", text)
    fixture = {
        "source": "jax/taylor_green_colab.ipynb stored outputs (cells 3-5)",
        "params": {"N": 128, "dt": 0.005, "reynolds": 1000, "total_time": 1.0, "save_every_time": 0.05},
        "times": sorted(ens),
        "enstrophy_str": [ens[t] for t in sorted(ens)],
        "enstrophy_digits": 7,
        "final_energy_str": final_energy,
        "max_divergence_str": max_div,
        "wall_time_s_str": run.group(3),
        "steps": int(run.group(2)),
    }
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "notebook_tgv_n128.json")
    with open(out, "w") as f:
        json.dump(fixture, f, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
