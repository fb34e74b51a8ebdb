"""Host logic of the autotuner on the CPU: the traverse_schedule word that
deepen_last_step builds, read the way the detector plans its BVTT steps
(leading steps from the word, 2 bits each, until a zero field; then the
uniform traverse_expand, which the tuner sets to expand + 1 with it)."""
from oibvh_amd.autotune import deepen_last_step


def plan(levels, word, uniform):
    """Levels each BVTT step descends until `levels` are covered (at least one
    step), as the detector's prepare does."""
    steps, done, lead = [], 0, True
    while not steps or done < levels:
        e = (word & 3) if lead else 0
        if e == 0:
            lead = False
            e = uniform
        word >>= 2
        steps.append(e)
        done += e
    return steps


def test_deepened_last_step_ends_exactly_at_the_leaves():
    seen = 0
    for depth in range(1, 26):
        for entry in range(depth):
            for expand in (1, 2, 3):
                w = deepen_last_step(depth, entry, expand)
                if not w:
                    continue
                seen += 1
                steps = plan(depth - entry, w, expand + 1)
                assert sum(steps) == depth - entry, (depth, entry, expand)
                assert steps[-1] == expand + 1 and all(s == expand for s in steps[:-1])
                # one front (one launch) fewer than uniform steps of `expand`
                assert len(steps) == len(plan(depth - entry, 0, expand)) - 1
    assert seen > 100


def test_no_word_where_no_exact_split_exists():
    assert deepen_last_step(20, 8, 3) == 0  # steps of 4 levels are not offered
    assert deepen_last_step(20, 17, 2) == 0  # 3 levels left: one step of 3 is the whole walk
    assert deepen_last_step(20, 8, 2) == 0  # 12 - 3 = 9 levels do not split into steps of 2
    assert deepen_last_step(20, 9, 2) != 0  # 11 - 3 = 8: four steps of 2, then one of 3
