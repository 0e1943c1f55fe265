"""Pins the oracle's restatement of the two pass controllers
(RTSTreeControllerT, OPX7UtilsPathGuiding.h:1794-1904; RTMortonQuadTreeControllerT,
OPX7UtilsMorton.h:780-905) against pass lists worked out by hand from their
rules:

- pass k asks for 2^k launches of spl samples, from a remaining budget rounded
  up to whole launches;
- the pass takes all that remains when the remainder after it would be less
  than twice the pass, or when the samples spent reach ratio x budget;
- state: Record until iteration > iterationForBuilt, then RecordAndSample, and
  Sample for the last pass; the Morton controller has Locate while iteration
  <= iterationForLocate;
- the controller executes until more than the budget has been spent, so after
  the last pass it asks for one more launch whose pass holds no samples.
"""
import pytest

R, RS, S, L = 0, 1, 2, 3


@pytest.mark.parametrize("budget,ratio,spl,built,want", [
    # 1+2+...+64 = 127 spent, 129 remain: 129 - 128 < 2 * 128, so the 8th pass takes them
    (256, 0.5, 1, 0, [(R, 1), (RS, 2), (RS, 4), (RS, 8), (RS, 16), (RS, 32), (RS, 64), (S, 129)]),
    # 255 spent, 745 remain: 745 - 256 = 489 < 512, so the 9th pass takes 745
    (1000, 0.5, 1, 0, [(R, 1), (RS, 2), (RS, 4), (RS, 8), (RS, 16), (RS, 32), (RS, 64), (RS, 128), (S, 745)]),
    # ratio: 31 spent >= 0.3 x 100 before pass 6, which takes the remaining 69
    (100, 0.3, 1, 0, [(R, 1), (RS, 2), (RS, 4), (RS, 8), (RS, 16), (S, 69)]),
    # 4 samples per launch: passes of 4, 8, 16, then 72 (72 - 32 = 40 < 64)
    (100, 0.5, 4, 0, [(R, 4), (RS, 8), (RS, 16), (S, 72)]),
    # budget 99 rounds up to 100 = 25 launches of 4
    (99, 0.5, 4, 0, [(R, 4), (RS, 8), (RS, 16), (S, 72)]),
    # trees sampled only after iteration 2
    (256, 0.5, 1, 2, [(R, 1), (R, 2), (R, 4), (RS, 8), (RS, 16), (RS, 32), (RS, 64), (S, 129)]),
])
def test_sdtree_controller_passes(orc, budget, ratio, spl, built, want):
    assert orc.ctrl_passes(0, budget, built, 0, ratio, spl) == want


def test_sdtree_controller_launch_events(orc):
    sch = orc.ctrl_schedule(0, 256, 0, 0, 0.5, 1)
    # 256 launches in the passes, then the extra launch (pass of 0 samples,
    # state Sample) after which 257 > 256 stops the controller
    assert len(sch) == 257 and sch[-1][:3] == (S, 8, 0)
    starts = [k for k, (_, _, _, ev) in enumerate(sch) if ev & 1]
    ends = [k for k, (_, _, _, ev) in enumerate(sch) if ev & 2]
    # Reset before the first launch of every pass but the last; Build at the end of every pass
    assert starts == [0, 1, 3, 7, 15, 31, 63]
    assert ends == [0, 2, 6, 14, 30, 62, 126, 255, 256]
    # budget 99 with 4 samples per launch: 25 launches spend 100 > 99, no extra launch
    sch = orc.ctrl_schedule(0, 99, 0, 0, 0.5, 4)
    assert len(sch) == 25 and sch[-1][2] == 72


@pytest.mark.parametrize("budget,built,locate,want", [
    # the bench's hash-grid schedule: Locate in pass 0, trees built from iteration 3
    (64, 3, 0, [(L, 1), (R, 2), (R, 4), (R, 8), (RS, 16), (S, 33)]),
    (256, 3, 1, [(L, 1), (L, 2), (R, 4), (R, 8), (RS, 16), (RS, 32), (RS, 64), (S, 129)]),
    (256, 0, 0, [(L, 1), (RS, 2), (RS, 4), (RS, 8), (RS, 16), (RS, 32), (RS, 64), (S, 129)]),
])
def test_morton_controller_passes(orc, budget, built, locate, want):
    assert orc.ctrl_passes(1, budget, built, locate, 0.5, 1) == want


def test_morton_controller_updates(orc):
    sch = orc.ctrl_schedule(1, 64, 3, 0, 0.5, 1)
    ends = [k for k, (_, _, _, ev) in enumerate(sch) if ev & 2]
    updates = [k for k, (_, _, _, ev) in enumerate(sch) if ev & 4]
    snaps = [k for k, (_, _, _, ev) in enumerate(sch) if ev & 8]
    # passes end at launches 0, 2, 6, 14, 30, 63 and the extra launch 64; the
    # building and sampling pyramids swap once iteration > 3, i.e. after the
    # 4th pass and every later one (the extra launch included); the
    # application snapshots the hash grid after every pass
    assert ends == [0, 2, 6, 14, 30, 63, 64]
    assert updates == [14, 30, 63, 64]
    assert snaps == ends
