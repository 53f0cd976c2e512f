"""The results pipeline's line contract (test infrastructure).

tests/d.make_tabular in the reference (d.make_tabular:27-44) turns a
program's stdout into one table row per observation: each line is matched
by `(.) ([^=]+)=?\\s*([^ ]*)\\s*(.*)`; tag A names the algorithm (spaces to
underscores), tag h records the text before '=' as `hash`, tags t, H and m
make a column (':' dropped, spaces to underscores) holding the first
whitespace-free token after '=', and tags p, C, R and # are skipped.  Any
other tag is reported as unrecognised.  `row` applies that contract to one
observation so that tests can check which columns the CLI's output yields.
"""
import re

LINE = re.compile(r"(.) ([^=]+)=?\s*([^ ]*)\s*(.*)")
SKIPPED = ("p", "C", "R", "#")
COLUMNS = ("t", "H", "m")


def row(lines):
    """Columns of one observation (the lines of one program run)."""
    datum = {}
    for line in lines:
        if not line.strip():
            continue
        m = LINE.match(line)
        if m is None:
            raise ValueError("bad line: %r" % line)
        tag, col, val = m.group(1).strip(), m.group(2).strip(), m.group(3).strip()
        if tag == "A":
            datum["Algorithm"] = col.replace(" ", "_")
        elif tag == "h":
            datum["hash"] = col
        elif tag in SKIPPED:
            continue
        elif tag in COLUMNS:
            datum[col.replace(":", "").replace(" ", "_")] = val
        else:
            raise ValueError("unrecognised tag %r in %r" % (tag, line))
    return datum


TIMER_COLUMNS = ["Step1_Initialize", "Step2_DetermineReceivers", "Step3_DetermineDonors",
                 "Step4_GenerateOrder", "Step5_FlowAcc", "Step6_Uplift", "Step7_Erosion",
                 "Overall", "Total_calculation_time"]
METADATA_COLUMNS = ["Random_seed", "Preset", "cell_steps_per_s", "step_cell_steps_per_s",
                    "roofline_bytes_per_cell_step", "hbm_roofline_frac", "nlevel_last_step",
                    "deepest_level_last_step", "tile_path_steps", "level_path_steps",
                    "tile_halo32_steps"]


def check_cli_row(datum, steps_run):
    """The columns the drop-in CLI must produce, with consistent values."""
    assert datum["Algorithm"] == "FastScape_RB_MI355X", datum
    assert "hash" in datum
    for c in TIMER_COLUMNS + METADATA_COLUMNS:
        assert c in datum, c
    timers = {c: int(datum[c]) for c in TIMER_COLUMNS}
    # every stage of the reference is timed, the tile path's fused ones too
    assert all(timers[c] > 0 for c in TIMER_COLUMNS[1:7]), timers
    assert sum(timers[c] for c in TIMER_COLUMNS[1:7]) <= timers["Overall"]
    assert int(datum["tile_path_steps"]) + int(datum["level_path_steps"]) == steps_run
    assert float(datum["roofline_bytes_per_cell_step"]) == 79.0
    frac = float(datum["hbm_roofline_frac"])
    rate = float(datum["step_cell_steps_per_s"])
    # both printed to 6 significant digits
    assert 0 < frac < 1 and abs(frac - rate * 79 / 8e12) <= 2e-5 * frac
    assert int(datum["deepest_level_last_step"]) == int(datum["nlevel_last_step"]) - 3
