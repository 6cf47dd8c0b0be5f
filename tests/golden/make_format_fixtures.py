"""Copies the reference's own input files used by the on-disk format tests into this
directory (data fixtures; gzip for the larger ones):
  test/1D/Sod79/INPUT/ic_x1.tec, bc_x1.txt           -> sod79_ic_x1.tec, sod79_bc_x1.txt
  test/2D/euler/prandtl-meyer/INPUT/ic.tec, bc.txt   -> prandtl_meyer_ic.tec.gz, prandtl_meyer_bc.txt.gz
  test/1D/Ein91/INPUT/ic_x8.tec, bc_x8.txt           -> ein91_ic_x8.tec.gz, ein91_bc_x8.txt.gz
  test/1D/Noh87/INPUT/ic_x8.tec, bc_x8.txt           -> noh87_ic_x8.tec.gz, noh87_bc_x8.txt.gz
  test/1D/Tor99/INPUT/ic_x8.tec, bc_x8.txt           -> tor99_ic_x8.tec.gz, tor99_bc_x8.txt.gz
  test/2D/euler/ramp-channel/INPUT/ic.tec, bc.txt    -> ramp_channel_ic.tec.gz, ramp_channel_bc.txt.gz
  test/2D/euler/oblique-shock/INPUT/ic.tec, bc.txt   -> oblique_shock_ic.tec.gz, oblique_shock_bc.txt.gz
Run in a container that holds the reference at /root/reference."""
import gzip
import os
import shutil

REF = "/root/reference/test"
HERE = os.path.dirname(os.path.abspath(__file__))


def _gzip(src, dst):
    """Byte-reproducible gzip (no file name or time stamp in the header)."""
    with open(src, "rb") as f, open(os.path.join(HERE, dst), "wb") as raw, \
            gzip.GzipFile(filename="", mode="wb", compresslevel=9, fileobj=raw, mtime=0) as g:
        g.write(f.read())


def main():
    shutil.copyfile(os.path.join(REF, "1D/Sod79/INPUT/ic_x1.tec"), os.path.join(HERE, "sod79_ic_x1.tec"))
    shutil.copyfile(os.path.join(REF, "1D/Sod79/INPUT/bc_x1.txt"), os.path.join(HERE, "sod79_bc_x1.txt"))
    for src, dst in (("ic.tec", "prandtl_meyer_ic.tec.gz"), ("bc.txt", "prandtl_meyer_bc.txt.gz")):
        _gzip(os.path.join(REF, "2D/euler/prandtl-meyer/INPUT", src), dst)
    for case in ("Ein91", "Noh87", "Tor99"):
        for src in ("ic_x8.tec", "bc_x8.txt"):
            _gzip(os.path.join(REF, "1D", case, "INPUT", src), "%s_%s.gz" % (case.lower(), src))
    for case in ("ramp-channel", "oblique-shock"):
        for src in ("ic.tec", "bc.txt"):
            _gzip(os.path.join(REF, "2D/euler", case, "INPUT", src), "%s_%s.gz" % (case.replace("-", "_"), src))


if __name__ == "__main__":
    main()
