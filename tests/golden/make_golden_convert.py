"""Writes the matching-format converter fixtures under tests/golden/convert/
from the reference's own transform-solution and convert-VNC-matching,
compiled from /root/reference into oracle/_ref (make -C oracle).

Each case is a directory holding the input file (its name matters: the
tools derive the output name from it), the file the reference tool wrote,
its stdout and its exit status (rc.txt).

Run in the build container only:  python tests/golden/make_golden_convert.py
"""
import os
import shutil
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")

# (case, tool, input file name, input bytes)
CASES = [
    ("generic_ok", "transform-solution", "solution.csv",
     b"Male ID,Female ID\n12,34\n1,2\r\n7,77\n100000,5\n"),
    ("generic_no_ext", "transform-solution", "solution",
     b"male,female\n3,4\n4,3\n"),
    ("generic_no_trailing_newline", "transform-solution", "sol.v2.csv",
     b"Male ID,Female ID\n5,6\n6,5"),
    ("generic_bad_id", "transform-solution", "bad.csv",
     b"Male ID,Female ID\n1,2\n3,x4\n5,6\n"),
    ("generic_blank_line", "transform-solution", "blank.csv",
     b"Male ID,Female ID\n1,2\n\n3,4\n"),
    ("prefixed_ok", "convert-VNC-matching", "matching.csv",
     b"Male Node ID,Female Node ID\nm12,f34\nm1,f2\r\nm7,f77\n"),
    ("prefixed_header_only", "convert-VNC-matching", "empty.csv",
     b"Male Node ID,Female Node ID\n"),
    ("prefixed_missing_prefix", "convert-VNC-matching", "mixed.csv",
     b"Male Node ID,Female Node ID\nm1,f2\n3,f4\nm5,f6\n"),
]


def run(tool_path, workdir, name):
    p = subprocess.run([tool_path, name], cwd=workdir, capture_output=True)
    return p.returncode, p.stdout


def output_name(tool, name):
    dot = name.rfind(".")
    stem = name[:dot] if dot >= 0 else name
    return stem + ("_matching.csv" if tool == "transform-solution" else "_generic.csv")


def main():
    out_root = os.path.join(HERE, "convert")
    for case, tool, name, data in CASES:
        d = os.path.join(out_root, case)
        shutil.rmtree(d, ignore_errors=True)
        os.makedirs(d)
        with tempfile.TemporaryDirectory() as tmp:
            with open(os.path.join(tmp, name), "wb") as f:
                f.write(data)
            rc, stdout = run(os.path.join(REF, tool), tmp, name)
            produced = os.path.join(tmp, output_name(tool, name))
            with open(os.path.join(d, "input_" + name), "wb") as f:
                f.write(data)
            if os.path.exists(produced):
                shutil.copy(produced, os.path.join(d, "expected_output"))
        with open(os.path.join(d, "stdout.txt"), "wb") as f:
            f.write(stdout)
        with open(os.path.join(d, "rc.txt"), "w") as f:
            f.write(f"{tool}\n{rc}\n")
        print(case, rc)


if __name__ == "__main__":
    main()
