"""Comparison helpers for parity tests."""
import numpy as np


def rel_err(a, b) -> float:
    """max |a - b| / max |b| (field-wise relative error; 0 when both are zero)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    scale = float(np.max(np.abs(b))) if b.size else 0.0
    d = float(np.max(np.abs(a - b))) if b.size else 0.0
    if scale == 0.0:
        return d
    return d / scale


def assert_close(a, b, tol, what=""):
    e = rel_err(a, b)
    assert e <= tol, f"{what}: relative error {e:.3e} > {tol:.1e}"
    return e


def write_packed_hybrid_scene(folder, width=4.0, height=4.0, freq=100, r=0.0575, jitter=0.0,
                              seed=0):
    """A hybrid scene directory (hybrid.xml, dem.xml, mpm.xml) in folder: the
    column-collapse fixture's settings and continuum file, with the grains
    replaced by a hexagonal packing of touching circles of radius about r
    filling [0, width] x [0, height] against the wall and a floor, and the
    resampling period set to freq.  Returns the hybrid.xml path."""
    import re
    from pathlib import Path

    import numpy as np

    src = Path(__file__).resolve().parent / "golden" / "assets" / "column-collapse"
    folder = Path(folder)
    folder.mkdir(parents=True, exist_ok=True)
    rng = np.random.default_rng(seed)
    dx, dy = 2.0 * r, np.sqrt(3.0) * r
    bodies = []
    y, row = r, 0
    while y + r <= height:
        x = r + (r if row % 2 else 0.0)
        while x + r <= width:
            rr = r * (1.0 + jitter * rng.uniform(-1.0, 1.0))
            bodies.append((x, y, rr))
            x += dx
        y += dy
        row += 1
    header = [ln for ln in (src / "dem.xml").read_text().splitlines()
              if "<rigid_body" not in ln and "<geometry" not in ln
              and "</rigidbody2d_scene>" not in ln]
    lines = header + ['  <static_plane x="0 0" n="0 1"/>']
    lines += [f'  <geometry type="circle" r="{b[2]:.17g}"/>' for b in bodies]
    lines += [f'  <rigid_body x="{b[0]:.17g} {b[1]:.17g}" theta="0" v="0 0" omega="0" '
              f'rho="2.53" geo_idx="{k}" fixed="0"/>' for k, b in enumerate(bodies)]
    lines.append("</rigidbody2d_scene>")
    (folder / "dem.xml").write_text("\n".join(lines) + "\n")
    (folder / "mpm.xml").write_text((src / "mpm.xml").read_text())
    hyb = (src / "hybrid.xml").read_text()
    hyb = re.sub(r'avoidAVoidFreq="\d+"', f'avoidAVoidFreq="{int(freq)}"', hyb)
    (folder / "hybrid.xml").write_text(hyb)
    return folder / "hybrid.xml"
