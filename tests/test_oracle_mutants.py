"""Shows that the pins of tests/test_oracle_pins.py are discriminating: each
one-token mutant of an oracle arm below is compiled into its own library, and
at least one of the pins listed for it must fail on that library.

The mutants are plausible slips of a restatement of the reference (an off-by
one comparison, a wrong table entry, a dropped sign, a swapped index), each in
the arm a pin covers.  The mutated source is written to a temporary folder;
the oracle in oracle/ is never touched.
"""
import concurrent.futures
import os
import shutil
import subprocess
import tempfile

import pytest

import alf_clip_rdo_cases
import binding
import sao_decide_cases
import test_oracle_pins as pins

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")

# (id, original text, replacement, occurrence (0-based, among identical texts), pins that must catch it)
MUTANTS = [
    ("tc-table-entry", "13, 14, 15, 17, 19, 21, 24, 25,", "13, 14, 15, 17, 19, 21, 23, 25,", 0,
     ["pin_deblock_tables"]),
    ("qp-joint-index", "if (c->joint_cbcr2) return c->qp_c[2];", "if (c->joint_cbcr2) return c->qp_c[1];", 0,
     ["pin_chroma_qp"]),
    ("qp-legacy-clamp", "clampi(c->qp_y + pc->qp_offset, 0, 63)", "clampi(c->qp_y + pc->qp_offset, 0, 65)", 0,
     ["pin_chroma_qp"]),
    ("qp-component", "pc->c_idx == 0 ? 0 : pc->c_idx - 1;", "pc->c_idx == 0 ? 0 : pc->c_idx - 0;", 0,
     ["pin_chroma_qp"]),
    ("cls-hv-tie", "if (sv > sh) {", "if (sv >= sh) {", 0, ["pin_alf_classification_ladder"]),
    ("cls-d-tie", "if (sd0 > sd1) {", "if (sd0 >= sd1) {", 0, ["pin_alf_classification_ladder"]),
    ("cls-prefer-tie", "(int64_t)d1 * hv0 > (int64_t)hv1 * d0", "(int64_t)d1 * hv0 >= (int64_t)hv1 * d0", 0,
     ["pin_alf_classification_ladder"]),
    ("cls-dirs2-bound", "hvd1 * 2 > 9 * hvd0", "hvd1 * 2 >= 9 * hvd0", 0, ["pin_alf_classification_ladder"]),
    ("cls-dirs1-bound", "(hvd1 > 2 * hvd0 ? 1 : 0)", "(hvd1 >= 2 * hvd0 ? 1 : 0)", 0,
     ["pin_alf_classification_ladder"]),
    ("cls-transpose-table", "TRANSPOSE_TABLE[8] = {0, 1, 0, 2, 2, 3, 1, 3};",
     "TRANSPOSE_TABLE[8] = {0, 1, 0, 2, 2, 3, 3, 1};", 0, ["pin_alf_classification_ladder"]),
    ("cls-dir-class", "(((dir1 & 1) << 1) + dir_s) * 5", "(((dir2 & 1) << 1) + dir_s) * 5", 0,
     ["pin_alf_classification_ladder"]),
    ("cls-vartab-entry", "VAR_TAB[16] = {0, 1, 2, 2, 2, 2, 2, 3,", "VAR_TAB[16] = {0, 1, 2, 2, 2, 2, 3, 3,", 0,
     ["pin_alf_classification_ladder"]),
    ("cls-activity-shift", "clampl(prod >> (bd - 1), 0, 15)", "clampl(prod >> (bd - 2), 0, 15)", 0,
     ["pin_alf_classification_bit_depth", "pin_alf_classification_ladder"]),
    ("cls-vb-ac", "min_y = -2; max_y = 3; ac = 3;", "min_y = -2; max_y = 3; ac = 2;", 0,
     ["pin_alf_classification_virtual_boundary"]),
    ("cls-vb-window", "min_y = -2; max_y = 3; ac = 3;", "min_y = -2; max_y = 5; ac = 3;", 0,
     ["pin_alf_classification_virtual_boundary"]),
    ("cls-vb-carve", "(ph - y_ctb) > ctb - 4;", "(ph - y_ctb) > ctb - 8;", 0,
     ["pin_alf_classification_virtual_boundary"]),
    ("cls-pad-bottom", "(ctb - 4 - y4) < 5;", "(ctb - 4 - y4) < 4;", 0, ["pin_alf_classification_vb_padding"]),
    ("cls-pad-top", "int has_top = y4 >= ctb - 4;", "int has_top = y4 > ctb - 4;", 0,
     ["pin_alf_classification_vb_padding"]),
    ("alf-transpose-1", "{9, 4, 10, 8, 1, 5, 11, 7, 3, 0, 2, 6}", "{9, 4, 10, 8, 1, 5, 11, 7, 3, 0, 6, 2}", 0,
     ["pin_alf_luma_taps_by_transpose"]),
    ("alf-transpose-3", "{9, 8, 10, 4, 3, 7, 11, 5, 1, 0, 2, 6}", "{9, 8, 10, 3, 4, 7, 11, 5, 1, 0, 2, 6}", 0,
     ["pin_alf_luma_taps_by_transpose"]),
    ("alf-clip-not-transposed", "- curr, -c[idx[k]], c[idx[k]])", "- curr, -c[k], c[k])", 0,
     ["pin_alf_luma_taps_by_transpose"]),
    ("alf-tap1-sign", "(CL(1, 1, y2) + CL(1, -1, -y2))", "(CL(1, 1, y2) + CL(1, -1, y2))", 0,
     ["pin_alf_luma_taps_by_transpose", "pin_alf_luma_table45_rows"]),
    ("alf-tap9-offset", "(CL(9, 3, 0) + CL(9, -3, 0))", "(CL(9, 3, 0) + CL(9, -2, 0))", 0,
     ["pin_alf_luma_taps_by_transpose", "pin_alf_luma_table45_rows"]),
    ("alf-class-row", "uint32_t fidx = cl & 0xff;", "uint32_t fidx = cl & 0x0f;", 0,
     ["pin_alf_luma_taps_by_transpose"]),
    ("alf-table45-row", "if (y == ctb - 7 || y == ctb - 2) {", "if (y == ctb - 8 || y == ctb - 2) {", 0,
     ["pin_alf_luma_table45_rows"]),
    ("alf-table45-shift", "o[0] = 10; o[1] = 0; o[2] = 0; o[3] = 0;", "o[0] = 9; o[1] = 0; o[2] = 0; o[3] = 0;", 0,
     ["pin_alf_luma_table45_rows"]),
    ("alf-table46-row", "if (y == h - 1 || y == h - 4) {", "if (y == h - 1 || y == h - 5) {", 0,
     ["pin_alf_chroma_table46_rows"]),
    ("cc-tap3-x", "(smp(lpre, lstride, clip, xl - 1, yl + yp1) - centre)",
     "(smp(lpre, lstride, clip, xl + 1, yl + yp1) - centre)", 0, ["pin_cc_alf_taps"]),
    ("cc-tap6-y", "(smp(lpre, lstride, clip, xl, yl + yp2) - centre)",
     "(smp(lpre, lstride, clip, xl, yl + yp1) - centre)", 0, ["pin_cc_alf_taps"]),
    ("cc-clip-low", "clampi(asr(sum + 64, 7), -half, half - 1)", "clampi(asr(sum + 64, 7), -half + 1, half - 1)", 0,
     ["pin_cc_alf_clip"]),
    ("cc-table47-row", "if (yl == ctb - 6 || yl == ctb - 3) {", "if (yl == ctb - 8 || yl == ctb - 3) {", 0,
     ["pin_cc_alf_taps"]),
    # (The bound of applyAlfLineBufBoundary itself has no observable mutant in
    # 4:2:0: it is 0 only in a last CTB row at most CtbSizeY - 4 high, where the
    # rows Table 47 would shorten lie past the picture's last row and clamp to
    # the same sample either way.)
    ("cc-table47-zero-row", "if (yl == ctb - 5 || yl == ctb - 4) {", "if (yl == ctb - 5 || yl == ctb - 2) {", 0,
     ["pin_cc_alf_taps"]),
    ("sao-tile-bottom", "ny1 >= clip[3];", "ny1 > clip[3];", 0, ["pin_sao_tile_edges"]),
    ("sao-tile-chroma-x0", "cc_[0] = cy_[0] / (int32_t)sw;", "cc_[0] = cy_[0] / (int32_t)1;", 0,
     ["pin_sao_tile_edges"]),
    ("tile-span", "if (v >= bds[i] && v < bds[i + 1]) {", "if (v >= bds[i] && v <= bds[i + 1]) {", 0,
     ["pin_tile_clip_equals_separate_pictures", "pin_sao_tile_edges"]),
    ("alf-tile-chroma-x1", "cc_[2] = (cy_[2] + (int32_t)sw - 1) / (int32_t)sw;",
     "cc_[2] = (cy_[2] + (int32_t)sw + 1) / (int32_t)sw;", 1, ["pin_tile_clip_equals_separate_pictures"]),
    # SAO encoder decision (src/sao_enc.rs), caught by tests/sao_decide_cases.py
    ("sao-band-window", "for (int start = 0; start < 28; start++) {", "for (int start = 0; start < 29; start++) {", 0,
     ["sao:case_band_31_is_never_in_a_window"]),
    ("sao-round-floor", "int64_t opt = (dif[b] + cnt[b] / 2) / cnt[b];",
     "int64_t opt = (dif[b] + cnt[b] / 2) / cnt[b] - (((dif[b] + cnt[b] / 2) % cnt[b]) < 0);", 0,
     ["sao:case_offset_rounds_toward_zero", "sao:case_band_offset_closed_form"]),
    ("sao-eo-tie", "if (d < best) {\n      best = d;\n      eo = t;", "if (d <= best) {\n      best = d;\n      eo = t;", 0,
     ["sao:case_band_offset_closed_form"]),
    ("sao-eo-over-bo", "if (has_eo) *out = eo;\n  else if (has_bo) *out = bo;",
     "if (has_bo) *out = bo;\n  else if (has_eo) *out = eo;", 0, ["sao:case_edge_offset_class_and_counts"]),
    ("sao-cr-free", "sao_decide_ctb_cr(&src->cr, &rec->cr, cx, cy, cw, ch, bd, &p->cb, &p->cr)",
     "sao_decide_ctb_c(&src->cr, &rec->cr, cx, cy, cw, ch, bd, &p->cr)", 0, ["sao:case_cr_follows_cb"]),
    ("sao-chroma-ceil", "uint32_t cx = x0 / 2, cy = y0 / 2, cw = w / 2, ch = h / 2;",
     "uint32_t cx = x0 / 2, cy = y0 / 2, cw = (w + 1) / 2, ch = (h + 1) / 2;", 0,
     ["sao:case_chroma_region_rounds_down"]),
    ("sao-merge-lambda", "8 * sao_chroma_indep_bits(&in);", "4 * sao_chroma_indep_bits(&in);", 0,
     ["merge:merge_case_rate_weight"]),
    ("sao-merge-band-bits", "bits += c[k]->sao_type_idx == VVCLF_SAO_BAND ? 16 :",
     "bits += c[k]->sao_type_idx == VVCLF_SAO_BAND ? 10 :", 0, ["merge:merge_case_rate_weight"]),
    ("sao-merge-above-bits", "cctb, bd, &a->cb, &a->cr) + 8 * 2;", "cctb, bd, &a->cb, &a->cr) + 8 * 1;", 0,
     ["merge:merge_case_above_costs_two_bits"]),
    ("clip-rdo-strict", "        if (t < best) {", "        if (t <= best) {", 0,
     ["clip:case_no_change_keeps_zero", "clip:case_equal_trials_keep_first_index"]),
    ("clip-rdo-order", "for (uint8_t c = 1; c <= 3; c++) {", "for (uint8_t c = 3; c >= 1; c--) {", 0,
     ["clip:case_equal_trials_keep_first_index"]),
    ("clip-rdo-commit", "aps->luma_clip_idx[k][j] = best_clip;", "aps->luma_clip_idx[k][j] = 0;", 0,
     ["clip:case_result_replays"]),
    ("clip-rdo-running", "      running = best;", "      running = running;", 0, ["clip:case_result_replays"]),
]


def mutate(src, old, new, occurrence):
    pos, start = -1, 0
    for _ in range(occurrence + 1):
        pos = src.find(old, start)
        assert pos >= 0, old
        start = pos + 1
    return src[:pos] + new + src[pos + len(old):]


@pytest.fixture(scope="module")
def mutant_libs():
    with open(os.path.join(ORACLE_DIR, "vvclf_oracle.c")) as fh:
        src = fh.read()
    tmp = tempfile.mkdtemp(prefix="vvclf_mutants_")

    def build(m):
        name, old, new, occ, _ = m
        c_path = os.path.join(tmp, name + ".c")
        with open(c_path, "w") as fh:
            fh.write(mutate(src, old, new, occ))
        so = os.path.join(tmp, name + ".so")
        subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-I", ORACLE_DIR, "-o", so, c_path, "-lpthread"],
                       check=True, capture_output=True)
        return name, so

    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        libs = dict(ex.map(build, MUTANTS))
    yield libs
    shutil.rmtree(tmp, ignore_errors=True)


def test_every_mutation_site_exists():
    with open(os.path.join(ORACLE_DIR, "vvclf_oracle.c")) as fh:
        src = fh.read()
    for name, old, _new, occ, _ in MUTANTS:
        assert src.count(old) >= occ + 1, name


@pytest.mark.parametrize("mutant", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_pins_catch_mutant(mutant, mutant_libs, vv):
    name, _old, _new, _occ, catchers = mutant
    lib = binding.load(vv, path=mutant_libs[name])
    failures = []
    for pin in catchers:
        try:
            if pin.startswith("sao:"):
                sao_decide_cases.ALL_CASES[pin[4:]](vv, lib.oracle_sao_decide_picture)
            elif pin.startswith("clip:"):
                alf_clip_rdo_cases.ALL_CASES[pin[5:]](vv, lib, lib.oracle_alf_clip_rdo_luma)
            elif pin.startswith("merge:"):
                sao_decide_cases.ALL_MERGE_CASES[pin[6:]](vv, lib.oracle_sao_chroma_merge,
                                                          lib.oracle_sao_decide_picture)
            else:
                pins.ALL_PINS[pin](lib, vv)
        except AssertionError as e:
            failures.append((pin, str(e)[:80]))
    assert failures, f"no pin caught mutant {name}"
