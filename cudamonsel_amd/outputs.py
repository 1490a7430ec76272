"""Result files of a scan or image job.

- The reference's line-scan text (LinesOnLayers.cu:331-398): a header of
  run parameters, the column line "beamE (eV)  x(nm)  y (nm)  BSE yield  SE
  yield", one row per pixel as five numbers in C's "%f" form
  (std::to_string), the elapsed seconds last.  SE yield = SE count / N in
  single precision ((float)totalSE / nTrajectories, :370); BSE yield = the
  back-scattered fraction less the SE yield (:371).
- Per-pixel BSE and SE maps of a raster as .npy (float32, rows = y).
- A job's complete tallies as one .npz.
"""
import numpy as np

COLUMNS = "beamE (eV)\tx(nm)\ty (nm)\tBSE yield\tSE yield"


def scan_yields(back, se, per_pixel):
    """(BSE, SE) yields per pixel as LinesOnLayers computes them."""
    se_f = (np.asarray(se).astype(np.float32) / np.float32(per_pixel)).astype(np.float64)
    back_f = np.asarray(back, np.float64) / float(per_pixel)
    return back_f - se_f, se_f


def _f(v):
    """C's "%f" (std::to_string of a double)."""
    return "%f" % v


def write_scan_text(path, beam_eV, centers_m, back, se, per_pixel, header=None, sigma_nm=None, elapsed_s=None):
    """Writes the five-column text in the reference's order: trajectories
    per position, the descriptive lines of `header` (name -> value; the
    reference writes the line geometry there), beam energy and size, a blank
    line, the column names, the rows and the elapsed seconds."""
    bse, sey = scan_yields(back, se, per_pixel)
    c = np.asarray(centers_m, np.float64).reshape(-1, 3) * 1e9
    lines = [f"# Trajectories at each landing position: {int(per_pixel)}"]
    for k, v in (header or {}).items():
        lines.append(f"{k}: {v}")
    lines.append(f"Beam landing energies (eV): {_f(beam_eV)}")
    if sigma_nm is not None:
        lines.append(f"Beam size (standard deviation, in nm): {_f(sigma_nm)}")
    lines.append("")
    lines.append(COLUMNS)
    for i in range(c.shape[0]):
        lines.append(" ".join(_f(v) for v in (beam_eV, c[i, 0], c[i, 1], bse[i], sey[i])))
    if elapsed_s is not None:
        lines.append(repr(float(elapsed_s)))
    with open(path, "w") as f:
        f.write("\n".join(lines))


def read_scan_text(path):
    """Rows of a five-column scan text (this module's or the reference's own
    output.txt): dict of float arrays beamE, x_nm, y_nm, bse, se, plus the
    header lines and the trailing elapsed time when present."""
    rows, header, elapsed = [], [], None
    seen_columns = False
    for line in open(path).read().splitlines():
        p = line.split()
        if not seen_columns:
            if line.strip().startswith("beamE"):
                seen_columns = True
            else:
                header.append(line)
            continue
        if len(p) == 5:
            rows.append([float(x) for x in p])
        elif len(p) == 1:
            try:
                elapsed = float(p[0])
            except ValueError:
                pass
    a = np.array(rows, np.float64).reshape(-1, 5)
    return dict(beamE=a[:, 0], x_nm=a[:, 1], y_nm=a[:, 2], bse=a[:, 3], se=a[:, 4], header=header,
                elapsed_s=elapsed)


def raster_maps(back, se, per_pixel, side_x, side_y=None):
    """(BSE map, SE map) as float32 arrays of shape (side_y, side_x), for a
    raster whose pixels run along x fastest."""
    side_y = side_y or side_x
    bse, sey = scan_yields(back, se, per_pixel)
    return (bse.astype(np.float32).reshape(side_y, side_x), sey.astype(np.float32).reshape(side_y, side_x))


def save_maps(prefix, back, se, per_pixel, side_x, side_y=None):
    """Writes <prefix>_bse.npy and <prefix>_se.npy; returns the two paths."""
    b, s = raster_maps(back, se, per_pixel, side_x, side_y)
    np.save(prefix + "_bse.npy", b)
    np.save(prefix + "_se.npy", s)
    return prefix + "_bse.npy", prefix + "_se.npy"


TALLY_KEYS = ("back", "fwd", "se", "se_sq", "bse_sq", "back_energy", "back_energy_hist", "fwd_energy_hist",
              "elevation_hist", "azimuth_hist")


def save_tallies(path, result, centers_m, per_pixel, beam_eV):
    """A job's tallies (an Engine result dict) with its pixel centres, as .npz."""
    arrays = {k: np.asarray(result[k]) for k in TALLY_KEYS}
    arrays["centers_m"] = np.asarray(centers_m, np.float64)
    arrays["per_pixel"] = np.int64(per_pixel)
    arrays["beam_eV"] = np.float64(beam_eV)
    np.savez(path, **arrays)


def load_tallies(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}
