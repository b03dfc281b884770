"""The miniqmc-style driver's host logic (src/Drivers/miniqmc.cpp:150-175 options, 509-551 report)."""
import xml.etree.ElementTree as ET

import pytest

from miniqmc_amd import driver


def test_options_map_tilings_to_workloads():
    assert driver.parse_args(["-g", "1 1 2"]).system == "nio64"
    assert driver.parse_args([]).system == "nio32"
    assert driver.parse_args(["--system", "nio8-cubic", "-g", "3 3 3"]).system == "nio8-cubic"
    a = driver.parse_args(["-w", "4", "-n", "3", "-k", "8", "-x", "2.0", "-s", "5"])
    assert (a.w, a.n, a.k, a.x, a.s) == (4, 3, 8, 2.0, 5)
    # miniqmc.cpp:150-175 defaults: no three-body Jastrow, one substep, fine timers
    assert (a.j, a.N, a.t) == (False, 1, "fine")
    a = driver.parse_args(["-j", "-N", "2", "-t", "coarse", "-v"])
    assert (a.j, a.N, a.t, a.v) == (True, 2, "coarse", True)


def test_version_exits_before_any_work(capsys):
    assert driver.main(["-V"]) == 0
    assert "miniqmc_amd" in capsys.readouterr().out


@pytest.mark.parametrize("argv", [["-g", "2 2 2"], ["-g", "1 1"], ["-g", "a b c"], ["-r", "0.4"], ["-w", "0"],
                                  ["-N", "0"], ["-t", "medium"], ["-a", "16"], ["-m", "2"]])
def test_options_reject_unsupported(argv):
    with pytest.raises(SystemExit):
        driver.parse_args(argv)


def test_throughput_formulas():
    timers = {"Total": (2.0, 1), "Diffusion": (1.0, 5), "Pseudopotential": (0.5, 5)}
    tp = driver.throughput(4, 10, timers)
    assert tp["total"] == pytest.approx(4 * 1000 / 2.0)
    assert tp["diffusion"] == pytest.approx(4 * 1000 / 1.0)
    assert tp["pseudopotential"] == pytest.approx(4 * 100 / 0.5)


def test_timer_table_and_info_xml(tmp_path):
    timers = {"Total": (2.0, 1), "Diffusion": (1.5, 3), "New Gradient": (0.5, 6)}
    table = driver.timer_table(timers).splitlines()
    # the reference's flat profile: header, then name (40 columns), time, calls, per call, per thread
    assert table[1] == "Flat profile" and len(table) == 5
    assert table[2][:40].strip() == "Diffusion" and table[2].split()[1:3] == ["1.5000", "3"]
    assert table[3][:40].strip() == "New Gradient" and table[3].endswith(" TIMER")
    assert set(driver.STAGE_TIMERS.values()) <= {"New Gradient", "Update", "Complete Updates", "Kinetic Energy"}
    timers = {"Total": (2.0, 1), "Diffusion": (1.5, 3)}
    path = tmp_path / "info_1_1_2.xml"
    driver.info_xml(str(path), 768, 3, timers)
    root = ET.parse(path).getroot()
    assert root.tag == "resources"
    assert [c.tag for c in root] == ["hardware", "timing", "particles", "run"]
    assert root.find("particles/particle/size").text == "768"
    assert root.find("run/driver/name").text == "miniqmc"
    assert root.find("run/driver/steps").text == "3"
    timing = root.find("timing")
    assert timing.find("max_stack_level_exceeded").text == "no"
    assert [t.find("name").text for t in timing.findall("timer")] == ["Total"]
    assert [t.find("name").text for t in timing.findall("timer/includes/timer")] == ["Diffusion"]


def test_stack_profile_nesting_and_xml_tree(tmp_path):
    """The reference's stack profile (NewTimer.cpp:281-331) and <timing> tree (NewTimer.cpp:333-395):
    scopes nested as they open in miniqmc.cpp:330-500, exclusive = inclusive less the children's."""
    timers = {"Setup": (1.0, 1), "Total": (10.0, 1), "Initialization": (0.5, 1), "Diffusion": (6.0, 5),
              "New Gradient": (2.0, 10), "Update": (3.0, 10), "Kinetic Energy": (0.25, 10),
              "Pseudopotential": (2.5, 5)}
    rows = driver.timer_stack(timers)
    assert [(lv, name) for lv, name, _, _, _ in rows] == [
        (0, "Setup"), (0, "Total"), (1, "Initialization"), (1, "Diffusion"), (2, "New Gradient"), (2, "Update"),
        (2, "Kinetic Energy"), (1, "Pseudopotential")]
    excl = {name: tx for _, name, _, tx, _ in rows}
    assert excl["Total"] == pytest.approx(10.0 - 0.5 - 6.0 - 2.5)
    assert excl["Diffusion"] == pytest.approx(6.0 - 2.0 - 3.0 - 0.25)
    assert excl["Update"] == pytest.approx(3.0)
    # device stages that overlap on several streams can sum past their parent: exclusive time 0
    over = dict(timers, Update=(5.0, 10))
    assert {n: tx for _, n, _, tx, _ in driver.timer_stack(over)}["Diffusion"] == 0.0
    lines = driver.stack_table(timers).splitlines()
    assert lines[0] == "Stack timer profile in seconds"
    assert lines[1].split() == ["Timer", "Inclusive_time", "Exclusive_time", "Calls", "Time_per_call"]
    assert lines[7].startswith("    Update") and lines[7].split()[1:4] == ["3.0000", "3.0000", "10"]
    assert lines[4].startswith("  Initialization")
    path = tmp_path / "info_1_1_1.xml"
    driver.info_xml(str(path), 192, 5, timers)
    timing = ET.parse(path).getroot().find("timing")
    assert [t.find("name").text for t in timing.findall("timer")] == ["Setup", "Total"]
    assert [t.find("name").text for t in timing.findall("timer/includes/timer")] == [
        "Initialization", "Diffusion", "Pseudopotential"]
    diff = timing.findall("timer/includes/timer")[1]
    assert [t.find("name").text for t in diff.findall("includes/timer")] == ["New Gradient", "Update", "Kinetic Energy"]
    assert float(diff.find("time_excl").text) == pytest.approx(0.75)
    assert timing.find("timer").find("includes") is None   # Setup has no children
