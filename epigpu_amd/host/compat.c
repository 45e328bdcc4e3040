/*
 * Level-1 drop-in for the reference's host entry point.
 *
 * perform_scan() keeps the signature of src/EpiOpenCL.c:1312 /
 * src/epiopencl.h:21 and reads the same run-configuration globals
 * (src/epiopencl.h:9-16, defaults src/main.c:244-249), so the reference's
 * analysis-mode main() links against this library in place of EpiOpenCL.c.
 * UCHUNKSIZE has no meaning here (one launch per block) and is ignored;
 * USAFE selects the safe-mode semantics. EPI_FORM selects the kernel form
 * ("plain" or "mfma", default mfma). Returns the total hit count as the
 * reference does, or a negative error code.
 */
#include <stdlib.h>
#include <string.h>

#include "episcan.h"
#include "epihost.h"

int UPERM, perm;
int UDEVICE;
int UCHUNKSIZE;
int UTEST;
int USAFE;
char plinkbin[100];
float UTHRESHOLD1, UTHRESHOLD2;
char UPHEN[1000];

int perform_scan(int *geno, chromosome *chrstat, ped *dat, map *genmap, int nid, int nsnp, int npack, int remain,
                 int nchr, char *filename)
{
    epi_run_opts o;
    epi_run_opts_default(&o);
    o.test = UTEST == 'i' ? 'i' : 'f';
    /* used as given, 0 included: the reference's main() sets 6.5 / 10.5 before parsing -F / -I
       (src/main.c:247-249) and its kernels take the globals unchanged (src/EpiOpenCL.c:1516-1517) */
    o.t1 = UTHRESHOLD1;
    o.t2 = UTHRESHOLD2;
    o.device = UDEVICE;
    o.perm = UPERM;
    o.safe = USAFE ? 1 : 0;
    o.datafile = plinkbin;
    o.phenfile = UPHEN;
    const char *form = getenv("EPI_FORM");
    if (form && strcmp(form, "plain") == 0) o.form = EPI_FORM_PLAIN;
    epi_run_report rep;
    int rc = epi_perform_scan(&o, geno, chrstat, dat, genmap, nid, nsnp, npack, remain, nchr, filename, &rep);
    if (rc) return rc;
    return (int)rep.total_hits;
}
