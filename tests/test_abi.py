"""C ABI and host layer of libctl on the CPU: exported symbols, the
input-folder parsers with the reference's error codes, and the compiled model
against the oracle's restatement of read_stoichiometry.  No kernel runs here."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, REPO


def test_exports_every_declared_symbol(ctl):
    hdr = open(os.path.join(REPO, "include", "ctl.h")).read()
    declared = set(re.findall(r"\b(ctl_\w+)\s*\(", hdr))
    assert declared == set(ctl.EXPORTS)
    L = ctl.lib()
    for name in declared:
        assert hasattr(L, name), name


def test_schlogl_folder_matches_oracle_compile(ctl, oracle_lib):
    m = ctl.Model.from_folder(os.path.join(GOLDEN, "schlogl"), 0, 1024)
    i = m.info()
    assert (i.S, i.R, i.K, i.SS, i.nL, i.nV, i.nF) == (3, 4, 100, 3, 5, 8, 4)
    assert i.x0_rows == 1 and i.c_rows == 1 and i.t_max == 10.0
    st = m.structure()
    o = oracle_lib.compile_model([[1, 2, 0], [0, 3, 0], [0, 0, 1], [0, 1, 0]],
                                 [[0, 3, 0], [1, 2, 0], [0, 1, 0], [0, 0, 1]], [1, 0, 1])
    assert st["hor"].tolist() == o["hor"].tolist() == [3, 3, 1]
    assert st["hort"].tolist() == o["hort"].tolist() == [1, 3, 1]


def test_random_network_compile_matches_oracle(ctl, oracle_lib):
    from cutauleaping_amd import networks
    net = networks.mass_action(30, 60, seed=3, check=False)
    feed = (np.arange(30) % 7 == 0).astype(np.int32)
    m = ctl.Model.from_arrays(net["left"], net["right"], net["x0"], net["c"], [1.0], feed=feed)
    st = m.structure()
    o = oracle_lib.compile_model(net["left"], net["right"], feed)
    assert np.array_equal(st["hor"], o["hor"]) and np.array_equal(st["hort"], o["hort"])
    lf = [(r, s, v) for r in range(60) for s, v in zip(o["Ls"][o["Lptr"][r]:o["Lptr"][r + 1]],
                                                        o["Lv"][o["Lptr"][r]:o["Lptr"][r + 1]])]
    assert st["left_flat"].tolist() == [list(t) for t in lf]
    ff = [(r, s, v) for r in range(60) for s, v in zip(o["Fs"][o["Fptr"][r]:o["Fptr"][r + 1]],
                                                        o["Fv"][o["Fptr"][r]:o["Fptr"][r + 1]])]
    assert st["fear_flat"].tolist() == [list(t) for t in ff]
    rf = [(r, s, v) for s in range(30) for r, v in zip(o["Vr"][o["Vptr"][s]:o["Vptr"][s + 1]],
                                                        o["Vv"][o["Vptr"][s]:o["Vptr"][s + 1]])]
    assert st["rav_flat"].tolist() == [list(t) for t in rf]


@pytest.fixture
def folder(tmp_path):
    d = tmp_path / "model"
    shutil.copytree(os.path.join(GOLDEN, "schlogl"), d)
    return d


def test_offset_rows_and_widths(ctl, folder):
    rows = [f"{100000}\t{250 + i}\t{200000}\n" for i in range(40)]
    (folder / "MX_0").write_text("".join(rows))
    (folder / "c_matrix").write_text("".join(f"3e-7\t1e-4\t1e-3\t{3.5 + i}\n" for i in range(40)))
    m = ctl.Model.from_folder(str(folder), offset=8, n=32)
    i = m.info()
    assert i.x0_rows == 32 and i.c_rows == 32 and i.row_offset == 8
    with pytest.raises(ctl.CtlError) as e:
        ctl.Model.from_folder(str(folder), offset=9, n=32)     # 9 + 32 > 40 rows
    assert e.value.code == ctl.CTL_E_ROWS


def test_error_codes(ctl, folder):
    (folder / "MX_0").write_text("100000\t250\n")
    with pytest.raises(ctl.CtlError) as e:
        ctl.Model.from_folder(str(folder), 0, 1)
    assert e.value.code == ctl.CTL_E_WIDTH
    os.remove(folder / "MX_0")
    with pytest.raises(ctl.CtlError) as e:
        ctl.Model.from_folder(str(folder), 0, 1)
    assert e.value.code == ctl.CTL_E_NOFILE
    (folder / "MX_0").write_text("100000\t250\t200000\n")
    (folder / "c_matrix").write_text("1\t2\t3\n")
    with pytest.raises(ctl.CtlError) as e:
        ctl.Model.from_folder(str(folder), 0, 1)
    assert e.value.code == ctl.CTL_E_ARGS
    (folder / "c_matrix").write_text("3e-7\t1e-4\t1e-3\t3.5\n")
    (folder / "right_side").write_text("0\t3\t0\n")
    with pytest.raises(ctl.CtlError) as e:
        ctl.Model.from_folder(str(folder), 0, 1)
    assert e.value.code == ctl.CTL_E_ARGS


def test_cs_vector_and_feed_parsing(ctl, folder):
    (folder / "cs_vector").write_text("1\n")
    m = ctl.Model.from_folder(str(folder), 0, 4)
    assert m.info().SS == 1
    os.remove(folder / "M_feed")
    m = ctl.Model.from_folder(str(folder), 0, 4)
    assert m.structure()["fear_flat"].shape[0] == 8    # nothing locked: A and B change too


def test_tts_vector_mismatch(ctl, folder):
    (folder / "p_rep").write_text("2\n")
    (folder / "ts_rep").write_text("1\n")
    (folder / "ts_numtgt").write_text("3\n")
    (folder / "tts_vector").write_text("100\n")
    m = ctl.Model.from_folder(str(folder), 0, 64 * 4)
    with pytest.raises(ctl.CtlError) as e:
        m.load_fitness(str(folder), 64, 4)             # 100 != 256 threads
    assert e.value.code == ctl.CTL_E_TTS


def test_no_device_fails_loudly(ctl):
    if ctl.device_count() > 0:
        pytest.skip("a GPU is visible")
    m = ctl.Model.from_folder(os.path.join(GOLDEN, "schlogl"), 0, 64)
    with pytest.raises(ctl.CtlError) as e:
        ctl.Sim(m, 64, 1)
    assert e.value.code == ctl.CTL_E_DEVICE


def test_offset_must_not_wrap_the_stream_index(ctl):
    """Philox stream indices are offset .. offset + n - 1 as 32-bit words: a
    negative offset or a range past 2^32 would alias another shard's streams,
    so both entry points reject it before touching the device."""
    m = ctl.Model.from_folder(os.path.join(GOLDEN, "schlogl"), 0, 64)
    for off in (-1, (1 << 32) - 63, 1 << 40):
        with pytest.raises(ctl.CtlError) as e:
            ctl.Sim(m, 64, 1, offset=off)
        assert e.value.code == ctl.CTL_E_ARGS
        with pytest.raises(ctl.CtlError) as e:
            ctl.Batch([m], [64], offset=off)
        assert e.value.code == ctl.CTL_E_ARGS


def test_cli_usage_exit_code(ctl):
    p = subprocess.run([ctl.CLI_PATH, os.path.join(GOLDEN, "schlogl"), "64"], capture_output=True, text=True)
    assert p.returncode == 255 and "please specify" in p.stdout


def test_feed_loader(ctl, tmp_path, capfd):
    """MX_feed: the species non-zero in the last line are fed; rows from the
    offset; width and row errors mirror MX_0's.  A missing MX_feed is the
    reference's warning, not an error (2phase_n-tau-leaping.cu:473-477)."""
    import shutil
    folder = tmp_path / "m"
    shutil.copytree(os.path.join(GOLDEN, "schlogl"), folder)
    m = ctl.Model.from_folder(str(folder), 0, 4)
    capfd.readouterr()
    m.load_feeds(str(folder), 0, 4)
    assert "WARNING: cannot read feed quantities" in capfd.readouterr().out
    (folder / "MX_feed").write_text("".join(f"0\t{10 + i}\t0\n" for i in range(4)))
    m.load_feeds(str(folder), 0, 4)
    with pytest.raises(ctl.CtlError) as e:
        m.load_feeds(str(folder), 2, 4)
    assert e.value.code == ctl.CTL_E_ROWS
    (folder / "MX_feed").write_text("0\t10\n0\t11\n")
    with pytest.raises(ctl.CtlError) as e:
        m.load_feeds(str(folder), 0, 2)
    assert e.value.code == ctl.CTL_E_WIDTH
    with pytest.raises(ctl.CtlError):
        m.set_feeds([5], [[1]])
    m.set_feeds([], np.zeros((1, 0), np.uint32))


def test_phase_clock_only_in_the_profiling_build(ctl):
    """ctl_phase_clock is exported by both builds; the normal library
    (libctl.so) has no phase clock and says so without touching a device,
    and the profiling build exports the same symbols."""
    import ctypes as C
    buf = np.zeros(16, np.uint64)
    assert ctl.lib().ctl_phase_clock(0, buf.ctypes.data, 16, 0) == ctl.CTL_E_STATE
    assert b"phase clock" in ctl.lib().ctl_last_error()
    prof = os.path.join(os.path.dirname(ctl.LIB_PATH), "libctl_phases.so")
    if not os.path.exists(prof):
        pytest.skip("profiling build not built")
    P = C.CDLL(prof)
    for name in ctl.EXPORTS:
        assert hasattr(P, name), name
