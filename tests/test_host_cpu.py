"""The gsi script host on the CPU: interpreter, expressions, field formats and the
error paths the reference scripts can hit before any GPU work starts
(src/input.cu:14-108, src/variable.cu:66-214, src/gs.cu:57-212)."""
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GSI = os.path.join(ROOT, "general-simulator_amd", "gsi")


def run(tmp_path, script, *args):
    p = tmp_path / "in.test"
    p.write_text(script)
    r = subprocess.run([GSI, str(p), *args], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    return r.returncode, r.stdout


def test_expressions_and_print(tmp_path):
    rc, out = run(tmp_path, ":steps=3000\nprint {steps/5} {2^3} {(1+2)*3}\n:x=steps\n:x+=1\nprint {x} {x>3000}\n"
                            "set gridsize 64 64 64\nprint {$gridsize}\n")
    assert rc == 0, out
    assert "600 8 9" in out and "3001 1" in out


def test_comments_semicolons_commas(tmp_path):
    rc, out = run(tmp_path, "print a # comment\n:y=2; print {y*y}\nprint 1,2,3\n")
    assert rc == 0
    lines = [l for l in out.splitlines() if not l.startswith(">>>")]
    assert "a" in lines and "4" in lines and "1 2 3" in lines


def test_readhere_writehere_and_text_format(tmp_path):
    script = ("readhere varianttensor\n3 2 3 3\n0.00, 0.0063, 0.0063 , 0.0063, 0.0, 0.0063 , 0.0063, 0.0063, 0.0\n"
              "1 2 3 4 5 6 7 8 9\nwrite varianttensor vt.data\n")
    rc, out = run(tmp_path, script)
    assert rc == 0, out
    lines = (tmp_path / "vt.data").read_text().split("\n")
    assert lines[0].split() == ["3", "2", "3", "3"]
    vals = [float(x) for x in lines[1:] if x.strip()]
    assert vals == [0.0, 0.0063, 0.0063, 0.0063, 0.0, 0.0063, 0.0063, 0.0063, 0.0, 1, 2, 3, 4, 5, 6, 7, 8, 9]


def test_binary_round_trip(tmp_path):
    # binary format: int32 rank, int32 dims, float32 values (include/datamain.th:356-418)
    vals = np.arange(6, dtype=np.float32) * 0.5
    with open(tmp_path / "f.bin", "wb") as f:
        f.write(struct.pack("iii", 2, 2, 3))
        f.write(vals.tobytes())
    rc, out = run(tmp_path, "set binary on\nread field f.bin\nwrite field g.bin\nset binary off\nwrite field g.txt\n")
    assert rc == 0, out
    raw = (tmp_path / "g.bin").read_bytes()
    assert struct.unpack("iii", raw[:12]) == (2, 2, 3)
    np.testing.assert_array_equal(np.frombuffer(raw[12:], np.float32), vals)
    assert (tmp_path / "g.txt").read_text().split("\n")[0].split() == ["2", "2", "3"]


def test_shell_and_backticks(tmp_path):
    rc, out = run(tmp_path, "shell echo hello\nprint `echo 7`\n")
    assert rc == 0 and "hello" in out and "\n7\n" in out


def test_errors_stop_the_script(tmp_path):
    rc, out = run(tmp_path, "sys nosuch\nprint never\n")
    assert rc == 1 and 'unknown system "nosuch"' in out and "never" not in out.replace(">>>print never", "")
    rc, out = run(tmp_path, "frobnicate 3\n")
    assert rc == 1
    rc, out = run(tmp_path, "run 10\n")
    assert rc == 1 and "system not set" in out
    rc, out = run(tmp_path, "read eta nofile.data\n")
    assert rc == 1 and "not found" in out
    rc, out = run(tmp_path, "print {undefined_var+1}\n")
    assert rc == 1


def test_missing_varianttensor_aborts_run(tmp_path):
    # the reference logs and carries on (src/gs.cu:180); here the failed Initialize stops the run
    rc, out = run(tmp_path, "sys mart\nset gridsize 16 16 16 1 1 1\nrun 10\n")
    assert rc == 1 and "strain tensor does not set" in out


def test_grid_validation(tmp_path):
    """Box grids (`gridsize nx ny nz dx dy dz` with unequal edges or spacings, src/dynamics_mart.cu:18)
    reach the library's box path; edges and spacings must be positive, and the stress plugin takes
    cubes only (both checked by the host before any GPU work)."""
    vt = "readhere varianttensor\n3 1 3 3\n1 0 0 0 1 0 0 0 1\n"
    rc, out = run(tmp_path, "sys mart\nset gridsize 16 0 32 1 1 1\n" + vt + "run 1\n")
    assert rc == 1 and "positive edges and spacings" in out
    rc, out = run(tmp_path, "sys mart\nset gridsize 16 16 32 1 -1 1\n" + vt + "run 1\n")
    assert rc == 1 and "positive edges and spacings" in out
    rc, out = run(tmp_path, "sys stress\nset gridsize 16 16 32 1 1 1\n" + vt + "run\n")
    assert rc == 1 and "stress plugin takes a periodic n x n x n grid" in out
    rc, out = run(tmp_path, "sys stress\nset gridsize 16 16 16 1 1 0.5\n" + vt + "run\n")
    assert rc == 1 and "stress plugin takes a periodic n x n x n grid" in out


def test_control_flow_loop_while_if(tmp_path):
    """loop / while / if-else-end with nesting (src/input.cu:254-336): the block runs again
    for every pass, `{}` is substituted per pass, `else` belongs to the innermost if."""
    script = ("set log off\n:s=0\nloop 4\n  :s=s+1\n  if (s>2)\n    print big{s}\n  else\n    print small{s}\n  end\nend\n"
              ":k=5\nwhile k>1\n  :k=k-2\n  loop 2\n    print k{k}\n  end\nend\n"
              "if (k<0)\n  print neg\nend\nprint done{s}{k}\n")
    rc, out = run(tmp_path, script)
    assert rc == 0, out
    lines = [l for l in out.splitlines() if not l.startswith(">>>") and not l.startswith("Input script")]
    assert lines == ["small1", "small2", "big3", "big4", "k3", "k3", "k1", "k1", "done41"]


def test_link_aliases_a_variable(tmp_path):
    # link target alias (src/gs.cu:110-115, include/datamap.th:126-139): the alias reads and
    # writes the target; an unknown target leaves everything as it was
    script = ("set deltatime 0.5\nset dt 7\nlink deltatime dt\nprint {$dt}\n"
              "set dt 0.25\nprint {$deltatime}\nlink nosuch alias2\nset alias2 3\nprint {$alias2} {$deltatime}\n")
    rc, out = run(tmp_path, script)
    assert rc == 0, out
    lines = [l for l in out.splitlines() if not l.startswith(">>>")]
    assert "0.5" in lines and "0.25" in lines and "3 0.25" in lines
    assert any("nosuch" in l for l in lines)


def test_control_flow_without_end_is_an_error(tmp_path):
    rc, out = run(tmp_path, "loop 3\nprint x\n")
    assert rc != 0 and "without end" in out


def read_bin(path):
    raw = open(path, "rb").read()
    rank = struct.unpack("i", raw[:4])[0]
    dims = struct.unpack("i" * rank, raw[4:4 + 4 * rank])
    return np.frombuffer(raw[4 + 4 * rank:], np.float32).reshape(dims)


def _sphere_count(n, ox, oy, oz, r):
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    a, b, d = i - ox, j - oy, k - oz
    inside = np.zeros((n, n, n), bool)
    for sa, sb, sd in [(0, 0, 0), (n, 0, 0), (-n, 0, 0), (0, n, 0), (0, -n, 0), (0, 0, n), (0, 0, -n)]:
        inside |= (a + sa) ** 2 + (b + sb) ** 2 + (d + sd) ** 2 <= r * r
    return inside


def test_cores_regular_places_spheres(tmp_path):
    """cores, regular method (src/dynamics_cores.cu:86-115): core cn at (n/2, n/2, n/C/2 + n/C cn),
    variant cn % V, eta = 1 and c = c2 inside (sphere plus one-period images), c = c1 outside."""
    n, C, r, V = 16, 2, 3, 3
    script = (f"sys cores\nset gridsize {n} {n} {n} 0.1 0.1 0.1\nset variantn {V}\nset coresn {C}\nset radius {r}\n"
              "set concentration 0.2 0.44\nset method regular\nset binary on\ndump out 1 eta concentration\nrun\n")
    rc, out = run(tmp_path, script)
    assert rc == 0, out
    eta = read_bin(tmp_path / "out" / "eta.1.data")
    c = read_bin(tmp_path / "out" / "concentration.1.data")
    assert eta.shape == (V, n, n, n)
    want_c = np.full((n, n, n), 0.2)
    for cn in range(C):
        ins = _sphere_count(n, n // 2, n // 2, n // C // 2 + n // C * cn, r)
        assert np.array_equal(eta[cn % V] > 0.5, ins) or cn % V != cn
        want_c[ins] = 0.44
    np.testing.assert_allclose(c, want_c.astype(np.float32), rtol=0, atol=0)


def test_cores_random_is_reproducible(tmp_path):
    script = ("sys cores\nset gridsize 16 16 16 0.1 0.1 0.1\nset variantn 4\nset coresn 3\nset radius 2\n"
              "set seed {s}\nset binary on\ndump o{s} 1 eta concentration\nrun\n")
    outs = []
    for s in (7, 7, 8):
        d = tmp_path / f"r{len(outs)}"
        d.mkdir()
        rc, out = run(d, script.replace("{s}", str(s)))
        assert rc == 0, out
        outs.append((read_bin(d / f"o{s}" / "eta.1.data"), read_bin(d / f"o{s}" / "concentration.1.data")))
    assert np.array_equal(outs[0][0], outs[1][0]) and not np.array_equal(outs[0][0], outs[2][0])
    eta, c = outs[0]
    inside = eta.sum(axis=0) > 0
    assert set(np.unique(eta)) <= {0.0, 1.0}
    assert np.all(c[inside] == np.float32(0.44)) and np.all(c[~inside] == np.float32(0.2))
    assert 0 < inside.sum() <= 3 * 33  # at most three spheres of radius 2 (33 lattice points each)


REF_SCRIPTS = "/root/reference/scripts"


def run_ref_script(tmp_path, name):
    path = os.path.join(REF_SCRIPTS, name)
    if not os.path.exists(path):
        pytest.skip(f"reference script {name} not present")
    r = subprocess.run([GSI, path], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    return [l for l in r.stdout.splitlines() if not l.startswith((">>>", "Input script"))]


def test_reference_script_if_loop_while(tmp_path):
    """scripts/in.if.loop.while unmodified: loop (a*2) runs 6 times counting a down from 3, while
    a>0 counts down to 0, and the if takes its else branch (a^2 = 4)."""
    out = run_ref_script(tmp_path, "in.if.loop.while")
    assert out == ["3", "2", "1", "0", "-1", "-2", "-3", "2", "1", "0", "0", "4"]


def test_reference_script_time_calculate(tmp_path):
    """scripts/in.time_calculate unmodified: backticks on `:x=` lines store the shell's output
    verbatim (all ten digits of `date +%s`), and {x2-x1} is the 3 s of `shell sleep 3`."""
    import re

    out = run_ref_script(tmp_path, "in.time_calculate")
    last = out[-1]
    m = re.match(r"since time point (\d{10}) to time point (\d{10}),?\s+(\d+) seconds have passed\.", last)
    assert m, out
    t1, t2, d = (int(g) for g in m.groups())
    assert t2 - t1 == d and d in (3, 4)


def test_reference_script_shell(tmp_path):
    out = run_ref_script(tmp_path, "in.shell")
    assert len(out) == 2 and all(l.strip().isdigit() for l in out)


def test_expression_text_is_kept_verbatim(tmp_path):
    """A variable assigned from a literal or a shell result keeps its text; operations print
    six significant digits, as the reference's stream formatting does."""
    rc, out = run(tmp_path, ":big=1234567891\nprint {big}\n:c=big\nprint {c} {big+1} {1/3}\n:s=`echo 0012`\nprint {s}\n")
    assert rc == 0, out
    assert "\n1234567891\n" in out and "1234567891 1.23457e+09 0.333333" in out and "\n0012\n" in out
