"""Copies the reference's own input files used by the on-disk format tests into this
directory (data fixtures; gzip for the larger ones):
  test/1D/Sod79/INPUT/ic_x1.tec, bc_x1.txt           -> sod79_ic_x1.tec, sod79_bc_x1.txt
  test/2D/euler/prandtl-meyer/INPUT/ic.tec, bc.txt   -> prandtl_meyer_ic.tec.gz, prandtl_meyer_bc.txt.gz
Run in a container that holds the reference at /root/reference."""
import gzip
import os
import shutil

REF = "/root/reference/test"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    shutil.copyfile(os.path.join(REF, "1D/Sod79/INPUT/ic_x1.tec"), os.path.join(HERE, "sod79_ic_x1.tec"))
    shutil.copyfile(os.path.join(REF, "1D/Sod79/INPUT/bc_x1.txt"), os.path.join(HERE, "sod79_bc_x1.txt"))
    for src, dst in (("ic.tec", "prandtl_meyer_ic.tec.gz"), ("bc.txt", "prandtl_meyer_bc.txt.gz")):
        with open(os.path.join(REF, "2D/euler/prandtl-meyer/INPUT", src), "rb") as f, \
                gzip.open(os.path.join(HERE, dst), "wb", compresslevel=9) as g:
            g.write(f.read())


if __name__ == "__main__":
    main()
