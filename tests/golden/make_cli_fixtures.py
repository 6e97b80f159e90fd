"""Writes the drop-in CLI fixtures of tests/golden/cli from the reference's
own, unchanged main (oracle/_ref/ref_harness with moments arguments; Main.c
is compiled from the reference sources in place by ``make -C oracle ref``).

    python tests/golden/make_cli_fixtures.py

Each case runs the reference binary on tests/golden/decks/squareSmall_cli.txt
(squareSmall with one output time, t = 0.05) with the case's flags, in a
scratch directory, and keeps:
    <case>/<output name>.hdr      the manifest (CREATED_UTC / GIT_COMMIT lines
                                  are free to differ)
    <case>/<output name>.dat.xz   the payload, xz-compressed
    <case>/stdout.txt, stderr.txt what the reference printed
    <case>/args.txt               the flags, one per line (paths relative to
                                  the repository root)
The cases exercise the flag paths of Main.c:70-106 and the restart reader
(Mesh.c:198-283): float-cast -Kn/-CFL/-R13 (tools.c:38-46), -double
(Float64 payload), -refine (Mesh.c:156-157), -limiter with a known and an
unknown name (Limiter.c:66-74), -sfile from the Float32 squareSmall golden.
"""
import lzma
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DECK = "tests/golden/decks/squareSmall_cli.txt"

CASES = {
    "flags_refine": ["-double", "-Kn", "0.01", "-CFL", "0.9", "-R13", "1", "-limiter", "superbee",
                     "-refine", "2"],
    "restart": ["-double", "-sfile", "tests/golden/squareSmall.ref"],
    "limiter_unknown": ["-limiter", "vanalbada", "-CFL", "0.5", "-Kn", "0.003"],
}


def main():
    sys.path.insert(0, ROOT)
    import oracle
    if not oracle.ref_available():
        sys.exit("oracle/_ref is not built (make -C oracle ref, needs the reference sources)")
    out_root = os.path.join(HERE, "cli")
    for case, flags in CASES.items():
        args = ["-bfile", DECK] + flags
        with tempfile.TemporaryDirectory() as tmp:
            absargs = [os.path.join(ROOT, a) if a.startswith("tests/") else a for a in args]
            r = subprocess.run([oracle.REF_HARNESS] + absargs + ["-out", tmp], capture_output=True,
                               text=True, cwd=tmp, timeout=600)
            if r.returncode != 0:
                sys.exit(f"{case}: reference exited {r.returncode}: {r.stderr}")
            dst = os.path.join(out_root, case)
            shutil.rmtree(dst, ignore_errors=True)
            os.makedirs(dst)
            for f in sorted(os.listdir(tmp)):
                src = os.path.join(tmp, f)
                if f.endswith(".hdr"):
                    shutil.copy(src, os.path.join(dst, f))
                elif f.endswith(".dat"):
                    with open(src, "rb") as fi, open(os.path.join(dst, f + ".xz"), "wb") as fo:
                        fo.write(lzma.compress(fi.read(), preset=9))
            # the scratch directory name appears in the log; keep it neutral
            with open(os.path.join(dst, "stdout.txt"), "w") as f:
                f.write(r.stdout.replace(tmp, "<out>").replace(ROOT, "<repo>"))
            with open(os.path.join(dst, "stderr.txt"), "w") as f:
                f.write(r.stderr.replace(tmp, "<out>"))
            with open(os.path.join(dst, "args.txt"), "w") as f:
                f.write("\n".join(args) + "\n")
        print(case, sorted(os.listdir(dst)))


if __name__ == "__main__":
    main()
