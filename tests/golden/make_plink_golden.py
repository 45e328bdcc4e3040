"""Writes the PLINK ingest fixtures under tests/golden/ from the reference's own
readplink.c (oracle/_ref/libreadplink.so, built by oracle/Makefile in the build
container where /root/reference is present).

  plink_small.ped / .map      60 individuals x 40 SNPs on 3 chromosomes (tests/plink_data.py,
                              seed 2024): missing calls written "0 0" and "0 A2", heterozygotes
                              in both allele orders, low call-rate SNPs and individuals, a
                              monomorphic SNP, non-positive positions, phenotypes -9
  plink_small.epigpu          the reference's -Dr output (dims, readped, readmap, pack,
                              createpackedbinary; src/readplink.c:1330-1345, :1420-1424)
  plink_small_cmn.epigpu      the reference's -Drcmn output (+ cleangeno, guessmissing and rnorm
                              on the stream of its unseeded rand())
  plink_small_cmn.x.ped/.map  the reference's -De of plink_small_cmn.epigpu
                              (extractfrombinary, src/readplink.c:150-224)
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle as O  # noqa: E402
from plink_data import make_dataset, write_ped_map  # noqa: E402


def main():
    ds = make_dataset(2024)
    ped, mp = os.path.join(HERE, "plink_small.ped"), os.path.join(HERE, "plink_small.map")
    write_ped_map(ds, ped, mp, zero_first=0.3)
    if O.ref_data_mode(ped, mp, os.path.join(HERE, "plink_small.epigpu")) is None:
        sys.exit("oracle/_ref/libreadplink.so is not built (make -C oracle)")
    cmn = os.path.join(HERE, "plink_small_cmn.epigpu")
    O.ref_data_mode(ped, mp, cmn, clean=True, impute=True, rnorm=True)
    O.ref_extract(cmn, os.path.join(HERE, "plink_small_cmn.x.ped"), os.path.join(HERE, "plink_small_cmn.x.map"))


if __name__ == "__main__":
    main()
