"""Mutation check of the oracle pins.

Each case plants one plausible transcription mistake in a scratch copy of
the oracle's C sources (never in oracle/ itself), rebuilds that copy with
gcc, and runs the closed-form pins against it: at least one pin must fail.
A mistake that no pin catches would be parity pinned by transcription
alone.  The unmutated copy must pass every pin (the control)."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PINS = ["test_oracle_pins.py", "test_oracle_momentum.py", "test_oracle_poisson.py", "test_oracle_grid.py"]

MOM, POI, GRID = "cv_oracle_momentum.c", "cv_oracle_poisson.c", "cv_oracle_grid.c"
MUTATIONS = [
    ("quick coefficient 1/6", MOM, "const double coef = 0.125;", "const double coef = 1.0 / 6.0;"),
    ("quick upwind sides swapped", MOM, "up = ucon + fabs(ucon), um = ucon - fabs(ucon);",
     "up = ucon - fabs(ucon), um = ucon + fabs(ucon);"),
    ("ff1 pressure term sign", MOM, "  r -= (cr->dpdc", "  r += (cr->dpdc"),
    ("ff1 central weight 1/2", MOM, "PP(o[0]+dn[0],o[1]+dn[1],o[2]+dn[2])) * 0.25;",
     "PP(o[0]+dn[0],o[1]+dn[1],o[2]+dn[2])) * 0.5;"),
    ("rk alpha2 and alpha3 swapped", MOM, "const double alfa[4] = {0.25, 1 / 3., 0.5, 1.};",
     "const double alfa[4] = {0.25, 0.5, 1 / 3., 1.};"),
    ("viscous transposed term sign", MOM, "fp->x += (g11 * dudc + g21 * dude + g31 * dudz + r11 * n0",
     "fp->x += (g11 * dudc + g21 * dude + g31 * dudz - r11 * n0"),
    ("viscous cross coefficient dropped", MOM, "fp->y += (g11 * dvdc + g21 * dvde + g31",
     "fp->y += (g11 * dvdc + 0 * dvde + g31"),
    ("viscous tangential weight 1/2", MOM, "D = &UU(dn[0], dn[1], dn[2]); f = 0.25;",
     "D = &UU(dn[0], dn[1], dn[2]); f = 0.5;"),
    ("rct basis average index", MOM, "0.5 * (V(eta,k,j,i).y + V(eta,k,j-1,i).y) * rc.y",
     "0.5 * (V(eta,k,j,i).y + V(eta,k,j,i).y) * rc.y"),
    ("poisson rhs without 1.5", POI, "/ c->st * COEF_TIME_ACCURACY;", "/ c->st * 1.0;"),
    ("projection without 1/1.5", POI, "* c->dt * c->st / COEF_TIME_ACCURACY;", "* c->dt * c->st;"),
    ("projection sign", POI, "if (d == 0) u->x -= corr; else if (d == 1) u->y -= corr;",
     "if (d == 0) u->x += corr; else if (d == 1) u->y -= corr;"),
    ("projection central weight 1/2", POI, "PP(o[0]+dn[0],o[1]+dn[1],o[2]+dn[2])) * 0.25;",
     "PP(o[0]+dn[0],o[1]+dn[1],o[2]+dn[2])) * 0.5;"),
    ("update pressure sign", POI, "S(p, k, j, i) += S(phi, k, j, i);", "S(p, k, j, i) -= S(phi, k, j, i);"),
    ("volumeflux correction halved", POI, "(flux + 0.0 /* FluxIntpSum */) / area : 0.;",
     "0.5 * (flux + 0.0 /* FluxIntpSum */) / area : 0.;"),
    ("divergence halved", POI, "V(uc,k,j,i).z - V(uc,k-1,j,i).z) * S(aj, k, j, i));",
     "V(uc,k,j,i).z - V(uc,k-1,j,i).z) * 0.5 * S(aj, k, j, i));"),
    ("divergence last index of a tie", POI, "if (dv > best) { best = dv;", "if (dv >= best) { best = dv;"),
    ("nvert restriction threshold", GRID, "if (prod > 0.1) S(nc,k,j,i)", "if (prod > 3.0) S(nc,k,j,i)"),
    ("nvert fill without the z test", GRID, "S(&tmp,k+1,j,i) + S(&tmp,k-1,j,i) > 1.1)", "1)"),
    ("face vector from an edge, not the midline", GRID,
     "dxde = 0.5 * (CX(co,k,j,i) + CX(co,k-1,j,i) - CX(co,k,j-1,i) - CX(co,k-1,j-1,i));",
     "dxde = (CX(co,k,j,i) - CX(co,k,j-1,i));"),
]


def _tree(dst):
    shutil.copytree(os.path.join(ROOT, "oracle"), dst / "oracle",
                    ignore=shutil.ignore_patterns("_build", "_ref", "__pycache__"))
    (dst / "curvib_old_amd").mkdir()
    for f in ("__init__.py", "cases.py"):
        shutil.copy(os.path.join(ROOT, "curvib_old_amd", f), dst / "curvib_old_amd" / f)
    (dst / "tests").mkdir()
    for f in PINS + ["conftest.py"]:
        shutil.copy(os.path.join(ROOT, "tests", f), dst / "tests" / f)


def _pins(w):
    env = dict(os.environ, OMP_NUM_THREADS="1")
    subprocess.run(["make", "-s", "-j4", "-C", str(w / "oracle")], check=True, capture_output=True, env=env)
    return subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider"] +
                          [os.path.join("tests", t) for t in PINS], cwd=w, capture_output=True, text=True, env=env)


@pytest.fixture(scope="module")
def base(tmp_path_factory):
    """A built, unmutated scratch copy; each mutant copies it (objects
    included) and rebuilds only the source it changes."""
    w = tmp_path_factory.mktemp("oracle_base")
    _tree(w)
    r = _pins(w)
    return w, r


def test_unmutated_copy_passes(base):
    assert base[1].returncode == 0, base[1].stdout[-2000:]


def _mutant(base_dir, n, fname, old, new):
    w = base_dir.parent / f"{base_dir.name}_mutant{n}"
    shutil.copytree(base_dir, w)
    path = w / "oracle" / fname
    text = path.read_text()
    assert old in text, f"mutation site missing in {fname}: {old}"
    path.write_text(text.replace(old, new))
    os.utime(path, None)
    try:
        return _pins(w)
    finally:
        shutil.rmtree(w, ignore_errors=True)


def test_every_mutation_is_caught(base):
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        runs = list(ex.map(lambda a: _mutant(base[0], *a),
                           [(n, f, o, w) for n, (_, f, o, w) in enumerate(MUTATIONS)]))
    missed = [m[0] for m, r in zip(MUTATIONS, runs) if not (r.returncode == 1 and "FAILED" in r.stdout)]
    assert not missed, f"mutations no pin catches: {missed}"
