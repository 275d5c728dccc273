"""Writes tests/golden/cli_contract.json: the status and message dl_run gives for a corpus of flag lists of both tools,
on a machine without a GPU (parse and prepare errors come before any GPU work; accepted flags stop at the device).

The messages are the command-line contract of reference src/cli.cpp (its flag grammar and error strings), which the C
ABI keeps; the fixture was taken from this library while its parser was still the line-by-line restatement of that
file, so that later parsers are held to the same contract.

    python tests/golden/make_cli_contract.py
"""
import ctypes as C
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "kit-sim_daylight_amd"))

DAYLIGHT = [
    [], ["-h"], ["--help"], ["-v"], ["--bogus"], ["--context"], ["--context", "x.obj"], ["x.obj"], ["--out", "a.txt"], ["--bench"],
    ["--bench", "10", "20"], ["--on", "faces"], ["--on", "edges"], ["--on"], ["--sides", "both"], ["--sides", "up"], ["--cull-backfaces"],
    ["--no-self-occlude"], ["--no-occlude", "a;b"], ["--no-occlude", "a;;b"], ["--no-occlude", ""], ["--north", "12.5"], ["--north", "abc"],
    ["--north", "inf"], ["--offset", "0.1"], ["--offset", ""], ["--virtual-ground", "off"], ["--virtual-ground", "maybe"],
    ["--optics", "a=0.2"], ["--optics", "a=0.2,0.5,glass"], ["--optics", "a=0.2,0.5,wood"], ["--optics", "=0.2"], ["--optics", "a"],
    ["--optics", "a/3=0.2,0.1"], ["--optics", "a/70000=0.2"], ["--optics", "*/3=0.2"], ["--optics", "/3=0.2"], ["--optics", "a=x"],
    ["--optics", "a=0.2,0.1,glass,x"], ["--optics", "a=1.5"], ["--optics", "a=0.5,0.6"], ["--optics", "a=0.1,0.5,glass"],
    ["--optics", "a=0,0.95,glass"], ["--optics", "a=0.2;b=0.3"], ["--optics", "Facade/East=0.2"], ["--optics", "a=0.2,"],
    ["--transmission", "on"], ["--transmission", "yes"], ["--mode", "ladybug"], ["--mode", "other"], ["--quantity", "radiation,sunhours"],
    ["--quantity", "radiation,foo"], ["--quantity", ""], ["--quantity", "illuminance"], ["--epw", "missing.epw"], ["--clear-sky"],
    ["--clear-sky", "--lat", "52"], ["--clear-sky", "--lat", "52", "--lon", "13"], ["--clear-sky", "--lat", "52", "--lon", "13", "--tz", "1"],
    ["--clear-sky", "--epw", "x.epw"], ["--lat", "abc", "--lon", "13", "--clear-sky"], ["--turbidity", "4", "--clear-sky", "--lat", "1", "--lon", "2"],
    ["--altitude", "100"], ["--tz"], ["--time", "06-21T12:00", "--lat", "52", "--lon", "13"], ["--time", "13-40T12:00", "--lat", "52", "--lon", "13"],
    ["--time", "06-21T12:00", "--from", "01-01T00:00"], ["--from", "01-01T00:00"], ["--to", "01-01T00:00"],
    ["--from", "01-01T00:00", "--to", "01-02T00:00", "--lat", "52", "--lon", "13"], ["--from", "x", "--to", "y"],
    ["--step", "0"], ["--step", "-5"], ["--step", "30", "--lat", "52", "--lon", "13"], ["--hours", "08:00-18:00"],
    ["--hours", "8:00-18:00", "--from", "01-01T00:00", "--to", "01-03T00:00", "--lat", "1", "--lon", "1"], ["--hours", "18:00-08:00"],
    ["--hours", "08:70-18:00"], ["--hours", "00:00-24:01"], ["--hours", "08:00"], ["--sky-density", "3", "--lat", "1", "--lon", "1"],
    ["--sky-density", "4"], ["--patch-samples", "2"], ["--ground", "0.3"], ["--fresnel", "off"], ["--fresnel", "x"],
    ["--bounces", "2"], ["--bounces", "x"], ["--bounces", "2.5"], ["--bounces", "30"], ["--bounce-samples", "10"],
    ["--bounce-samples", "65536"], ["--bounces", "2", "--mode", "ladybug"], ["--transmission", "on", "--mode", "ladybug"],
    ["--timeseries", "s.npy"], ["--dump-bits", "b.npy"], ["--dump-sky", "s.txt"], ["--dump-sun", "s.txt"], ["-5"],
    ["--lat", "52", "--lon", "13"], ["--lat", "52", "--lon", "13", "--quantity", "sunhours"], ["--lat", "52"],
    ["--clear-sky", "--lat", "52", "--lon", "13", "--quantity", "sunhours", "--time", "02-30T10:00"],
]
VIEWS = [
    [], ["-h"], ["--bogus"], ["--fov", "360"], ["--fov", "90,-10,10"], ["--fov", "90,-10"], ["--fov", "x"], ["--fov", "400"],
    ["--fov", "120,10,-10"], ["--fov", "120,0,0"], ["--fov", "120,0,0", "--distance", "isovist"], ["--cell", "0"], ["--cell", "x"],
    ["--cell", "1,2"], ["--cell", "0.01"], ["--jitter", "off"], ["--jitter", "x"], ["--distance", "mean,min"], ["--distance", "none"],
    ["--distance", "mean,foo"], ["--distance", "group"], ["--distance", "isovist", "--fov", "4", "--cell", "2"], ["--far", "0.05"],
    ["--far", "x"], ["--far", "500"], ["--ground-z", "1.5"], ["--ground-z", "inf"], ["--ground-z", ""], ["--reach", "en17037"],
    ["--reach", "6,20"], ["--reach", "6,x"], ["--reach", "1,2,3,4,5,6,7,8,9"], ["--reach", "6,6"], ["--reach", "-1"],
    ["--target", "a=park"], ["--target", "a=park;b=water"], ["--target", "a"], ["--target", "=park"], ["--target", "a=park;"],
    ["--target", ";a=park"], ["--target", "a="], ["--target", "a=sky"], ["--target", "a=b c"], ["--dump-rays", "r.npy"],
    ["--north", "x"], ["--offset", "1.6"], ["--no-occlude", "a"], ["--optics", "a=0.2,0.5,glass"], ["--optics", "a=2"],
    ["--transmission", "on"], ["--on", "vertices"], ["--sides", "back"], ["--context", "c.obj"], ["--out", "o.npy"], ["--bench"],
    ["--fov", "360", "--cell", "0.1"],
]


def main():
    import daylight_amd as d
    import numpy as np
    L = d.lib()
    m = L.dl_mesh_create(b"a", np.array([0, 0, 0, 1, 0, 0, 1, 1, 0], np.float64).ctypes.data, 3, np.array([3], np.int32).ctypes.data, 1,
                         np.array([0, 1, 2], np.int32).ctypes.data, 3, None)
    assert m
    hs = (C.c_void_p * 1)(m)
    out = {}
    for tool, corpus in ((0, DAYLIGHT), (1, VIEWS)):
        rows = []
        for argv in corpus:
            av = (C.c_char_p * max(1, len(argv)))(*[a.encode() for a in argv])
            res = C.c_void_p()
            rc = L.dl_run(tool, len(argv), av, hs, 1, None, 0, None, C.byref(res))
            msg = (L.dl_last_error() or b"").decode()
            if res.value:
                L.dl_result_free(res)
            rows.append({"argv": argv, "status": rc, "message": msg})
        out["daylight" if tool == 0 else "views"] = rows
    L.dl_mesh_free(m)
    with open(os.path.join(HERE, "cli_contract.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
