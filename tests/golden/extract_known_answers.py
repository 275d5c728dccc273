"""Writes tests/golden/known_answers.json: the known-answer data the reference's own tests and docs hold for this path.

Sources (read as text from the reference checkout; run once where it exists, the JSON is committed):
  tests/check_solar.cpp:18-56   Ladybug Sunpath goldens (lat, lon, tz, hoy, leap, alt, azi, up)
  tests/check_solar.cpp:80-85   sun vector at hoy 4000.5 Boston, and turned by north 30
  tests/check_solar.cpp:89-94   pvlib Ineichen goldens (z, TL, altitude, ghi, dni, dhi)
  tests/check_solar.cpp:101     Spencer extraterrestrial irradiance at doy 1, 172, 355
  tests/check_solar.cpp:112     Boston TL 3 clear-sky annual sums (kWh/m2)
  tests/check_sky.cpp:33-34     Ladybug patch-centre altitudes (degrees) for sky density 1
  docs/daylight.md:133-150      tests/box.obj, clear-sky Berlin, radiation + sunhours, as printed (%.10g)
  docs/daylight.md:163-180      tests/box.obj, sun hours of 21 June in 10-minute steps, as printed
  docs/views.md:333-350         tests/box.obj, views defaults (jitter on), as printed
Only numbers are taken; no source text is kept.
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def numbers(s):
    return [float(x) for x in re.findall(r"-?\d+(?:\.\d+)?(?:e[-+]?\d+)?", s)]


def block_after(md, marker):
    """The first fenced block after the line holding marker, as rows of numbers."""
    lines = md.splitlines()
    k = next(i for i, l in enumerate(lines) if marker in l)
    numeric = re.compile(r"^-?\d[\d.e+-]*( -?\d[\d.e+-]*)*$")
    start = next(i for i in range(k + 1, len(lines)) if numeric.match(lines[i].strip()))
    rows = []
    for l in lines[start:]:
        if not numeric.match(l.strip()):
            break
        rows.append([float(x) for x in l.split()])
    return rows


def main():
    solar = open(os.path.join(REF, "tests/check_solar.cpp")).read()
    sky = open(os.path.join(REF, "tests/check_sky.cpp")).read()
    out = {}
    gold = re.search(r"kGold\[\] = \{(.*?)\n\};", solar, re.S).group(1)
    out["ladybug_sun"] = [numbers(m) for m in re.findall(r"\{([^{}]*)\}", gold)]
    vec = re.search(r"sun_noaa\(42\.35, -71\.07, -5, 6, 16, 16\.5, false\);.*?CHECK\((.*?)\);\s*const float3 n = rotate_z\(s\.vec, 30\);\s*CHECK\((.*?)\);", solar, re.S)
    out["sun_vec_4000_5"] = numbers(vec.group(1))[0::2][:3]
    out["sun_vec_4000_5_north30"] = numbers(vec.group(2))[0::2][:2]
    cs = re.search(r"kCs\[\] = \{(.*?)\};", solar, re.S).group(1)
    out["ineichen"] = [numbers(m) for m in re.findall(r"\{([^{}]*)\}", cs)]
    sp = re.search(r"extraterrestrial_irradiance\(1\), ([\d.]+)\) && rel\(extraterrestrial_irradiance\(172\), ([\d.]+)\) && rel\(extraterrestrial_irradiance\(355\), ([\d.]+)\)", solar)
    out["spencer"] = {"1": float(sp.group(1)), "172": float(sp.group(2)), "355": float(sp.group(3))}
    bs = re.search(r"sdni / 1000 - ([\d.]+)\).*?sdhi / 1000 - ([\d.]+)\)", solar)
    out["boston_clear_sums"] = [float(bs.group(1)), float(bs.group(2))]
    out["alt_ladybug_mf1"] = numbers(re.search(r"altLB\[7\] = \{(.*?)\}", sky).group(1))
    day = open(os.path.join(REF, "docs/daylight.md")).read()
    out["box_berlin_radiation_sunhours"] = block_after(day, "daylight tests/box.obj --clear-sky --lat 52.52 --lon 13.41 --quantity radiation,sunhours")
    out["box_sunhours_june21"] = block_after(day, "daylight tests/box.obj --quantity sunhours --lat 52.52 --lon 13.41 --tz 1 --from 06-21T00:00")
    views = open(os.path.join(REF, "docs/views.md")).read()
    out["box_views_defaults"] = block_after(views, "**`tests/box.obj`, defaults**")
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump(out, f, indent=1)
    print({k: len(v) for k, v in out.items()})


if __name__ == "__main__":
    main()
