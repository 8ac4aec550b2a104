"""Writes viridis_lut.h: the 256-entry viridis colour map as 8-bit BGR
triplets, round-half-even of 255 x the published float table (matplotlib's
`viridis` ListedColormap data).  The header is committed; this script only
regenerates it (matplotlib is not needed at run time)."""
import os

import numpy as np
from matplotlib import colormaps


def main():
    rgb = np.asarray(colormaps["viridis"].colors, np.float64)
    assert rgb.shape == (256, 3)
    q = np.rint(rgb * 255.0).astype(np.uint8)          # numpy rint: half to even
    lines = ["// Generated by gen_viridis.py from the published viridis table; BGR order.",
             "#pragma once", "#include <cstdint>", "",
             "static const uint8_t kViridisBGR[256][3] = {"]
    for r, g, b in q:
        lines.append(f"    {{{b}, {g}, {r}}},")
    lines.append("};")
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "viridis_lut.h")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("wrote", out)


if __name__ == "__main__":
    main()
