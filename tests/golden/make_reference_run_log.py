"""Derive tests/golden/reference_run_log.json from the reference's committed
run log CudaMONSEL/CudaMONSEL/outputs/output_release.txt (a 0.5 keV PMMA
line scan, 100 primaries per pixel): the seed and the ten rand()/RAND_MAX
values it prints, the far-field (y > 700 nm, flat PMMA) mean yields, the
per-x mean yields over the line (y < 480 nm) and the elapsed time."""
import json
import sys

import numpy as np

src = sys.argv[1] if len(sys.argv) > 1 else \
    "/root/reference/CudaMONSEL/CudaMONSEL/outputs/output_release.txt"
lines = open(src).read().splitlines()
seed = int(lines[1].split(":")[1])
vals = [float(x) for x in lines[2:12]]
rows = []
for line in lines:
    p = line.split()
    if len(p) == 5:
        try:
            rows.append([float(x) for x in p])
        except ValueError:
            pass
a = np.array(rows)
far = a[a[:, 2] > 700]
on = a[a[:, 2] < 480]
prof = {}
for x in np.unique(on[:, 1]):
    m = on[:, 1] == x
    prof["%.1f" % x] = [float(on[m, 3].mean()), float(on[m, 4].mean()), int(m.sum())]
g = dict(source="CudaMONSEL/CudaMONSEL/outputs/output_release.txt",
         msvc_seed=seed, msvc_first_values=vals,
         flat_far={"beam_eV": 500.0, "primaries_per_row": 100, "rows": int(far.shape[0]),
                   "bse_mean": float(far[:, 3].mean()), "bse_sem": float(far[:, 3].std() / np.sqrt(len(far))),
                   "se_mean": float(far[:, 4].mean()), "se_sem": float(far[:, 4].std() / np.sqrt(len(far)))},
         line_profile_x_nm={"note": "mean BSE, SE over y<480 nm per x column (line h120 w80 3deg r20)",
                            "columns": prof},
         elapsed_s=float(lines[-1]))
json.dump(g, open(__file__.replace("make_reference_run_log.py", "reference_run_log.json"), "w"), indent=1)
# the head of the run log (header, column line, first rows) and its last
# line (elapsed seconds), for the parser of the five-column text
with open(__file__.replace("make_reference_run_log.py", "reference_output_excerpt.txt"), "w") as f:
    f.write("\n".join(lines[:30] + lines[-1:]))
