"""Copy the example input files of the reference (fullrmc Examples/) that
the realistic parity tests use into gzip fixtures next to this script.
These are the reference's own data files (structures and experimental
curves), compressed, not code.  Run from the repository root where the
reference tree is mounted at /root/reference."""
import gzip
import os
import shutil

SRC = "/root/reference/Examples"
FILES = {
    "niti_system.pdb.gz": "atomicNiTi/system.pdb",
    "niti_experimental.gr.gz": "atomicNiTi/experimental.gr",
    "niti_experimental.fq.gz": "atomicNiTi/experimental.fq",
    "removes_system.pdb.gz": "removes/system.pdb",
    "removes_pdf.exp.gz": "removes/pdf.exp",
}

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "examples")
    os.makedirs(out, exist_ok=True)
    for dst, src in FILES.items():
        with open(os.path.join(SRC, src), "rb") as fi, gzip.GzipFile(os.path.join(out, dst), "wb", 9,
                                                                     mtime=0) as fo:
            shutil.copyfileobj(fi, fo)
        print(dst, os.path.getsize(os.path.join(out, dst)))
