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


@pytest.mark.parametrize("argv", [["-g", "2 2 2"], ["-g", "1 1"], ["-g", "a b c"], ["-r", "0.4"], ["-w", "0"]])
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
    assert sorted(t.find("name").text for t in root.find("timing")) == ["Diffusion", "Total"]
