"""Shows that the nonlinear pins (_nonlinear_pins.py) catch plausible slips in
the oracle's restatement of src/solver/cpu_solver/cpu_solver.cpp:50-194.

Each case copies oracle/gpe_oracle.c into a temporary directory, plants one
mistake by an exact text substitution, builds the copy with gcc and runs every
pin on it: at least one pin must fail.  Everything runs on the CPU; the oracle
in the tree is not touched."""
import os
import subprocess

import pytest

import _nonlinear_pins as pins
from oracle import pyoracle

SRC = os.path.join(os.path.dirname(pyoracle.__file__), "gpe_oracle.c")

MUTATIONS = {
    # IT nonlinear step (cpu_solver.cpp:86-91)
    "it_contact_drops_cdd_over_3": (
        "s->p.imag_time_dt * ((s->p.ggp11 - s->p.cdd / 3) * nrm * c * w +",
        "s->p.imag_time_dt * ((s->p.ggp11) * nrm * c * w +"),
    "it_lhy_uses_n_not_n15": (
        "s->p.gamma * (ab * ab * ab) * c * s->p.w_15);",
        "s->p.gamma * (ab * ab * ab) * c * w);"),
    "it_contact_drops_n": (
        "* nrm * c * w +",
        "* nrm * c +"),
    # IT linear step (cpu_solver.cpp:60-69)
    "it_linear_drops_dipolar": (
        "c * (v + s->p.cdd * s->fi3d[IDX(s, i, j, k)]);",
        "c * v;"),
    "it_linear_dipolar_sign": (
        "c * (v + s->p.cdd * s->fi3d[IDX(s, i, j, k)]);",
        "c * (v - s->p.cdd * s->fi3d[IDX(s, i, j, k)]);"),
    # RT potential half step (cpu_solver.cpp:113-127)
    "rt_phase_sign": (
        "double sn = sin(-dt_factor * total), cs = cos(-dt_factor * total);",
        "double sn = sin(dt_factor * total), cs = cos(dt_factor * total);"),
    "rt_full_dt_not_half": (
        "const double dt_factor = s->p.real_time_dt / 2.;",
        "const double dt_factor = s->p.real_time_dt;"),
    "rt_contact_drops_cdd_over_3": (
        "double v_int = (s->p.ggp11 - s->p.cdd / 3) * density * w",
        "double v_int = s->p.ggp11 * density * w"),
    "rt_lhy_uses_n_not_n15": (
        "s->p.gamma * (ab * ab * ab) * s->p.w_15;",
        "s->p.gamma * (ab * ab * ab) * w;"),
    "rt_dipolar_sign": (
        "double total = s->pote[id] + s->p.cdd * s->fi3d[id] + v_int;",
        "double total = s->pote[id] - s->p.cdd * s->fi3d[id] + v_int;"),
    "rt_rotation_conjugated": (
        "s->cpsi[id] = (re * cs - im * sn) + I * (re * sn + im * cs);",
        "s->cpsi[id] = (re * cs + im * sn) + I * (-re * sn + im * cs);"),
    # energies (cpu_solver.cpp:164-189)
    "energy_bmf_n_power": (
        "e_bmf *= p->dxdydz * pow(p->n_atoms, 2.5);",
        "e_bmf *= p->dxdydz * pow(p->n_atoms, 2.0);"),
    "energy_int_drops_half": (
        "e_int += 0.5 * p->ggp11 * nrm * nrm;",
        "e_int += p->ggp11 * nrm * nrm;"),
    "energy_dd_local_part": (
        "e_ext -= p->cdd / 3. * pow(nrm, 2) / 2. * pow(p->n_atoms, 2);",
        "e_ext -= p->cdd / 3. * pow(nrm, 2) * pow(p->n_atoms, 2);"),
}


def build_copy(tmp_path, old, new):
    text = open(SRC).read()
    assert text.count(old) == 1, f"mutation target not unique in gpe_oracle.c: {old!r}"
    src = tmp_path / "gpe_oracle.c"
    src.write_text(text.replace(old, new))
    hdr = os.path.join(os.path.dirname(SRC), "gpe_oracle.h")
    lib = tmp_path / "libgpe_oracle_mut.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-fopenmp", "-std=gnu11", f"-I{os.path.dirname(hdr)}", "-shared",
                    "-o", str(lib), str(src), "-lm"], check=True)
    return pyoracle.bind(str(lib))


def failing_pins(L):
    failed = []
    for pin in pins.ALL_PINS:
        try:
            pin(L)
        except AssertionError as e:
            failed.append(f"{pin.__name__}: {e}")
    return failed


def test_unmutated_copy_passes(tmp_path):
    L = build_copy(tmp_path, "void orc_calc_norm(orc_solver *s) {", "void orc_calc_norm(orc_solver *s) {")
    assert failing_pins(L) == []


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_mutation_is_caught(tmp_path, name):
    old, new = MUTATIONS[name]
    L = build_copy(tmp_path, old, new)
    failed = failing_pins(L)
    print(name, "caught by", failed)
    assert failed, f"mutation {name} passes every nonlinear pin"
