/* offsetof() of the C declarations in include/champ_gpu.h, printed in the
 * same "name offset" form as layout_check.f90. */
#include <stddef.h>
#include <stdio.h>
#include "champ_gpu.h"

#define P(f) printf("%s %zu\n", #f, offsetof(champ_system_t, f))
#define Q(f) printf("stats.%s %zu\n", #f, offsetof(champ_sweep_stats, f))
#define V(f) printf("vmc.%s %zu\n", #f, offsetof(champ_vmc_result, f))
#define D(f) printf("dmcstep.%s %zu\n", #f, offsetof(champ_dmc_step_stats, f))
#define R(f) printf("dmcres.%s %zu\n", #f, offsetof(champ_dmc_result, f))

int main(void)
{
    P(nelec); P(nup); P(ncent); P(nctype); P(cent); P(iwctype); P(znuc); P(nr); P(grid_r0); P(grid_arg);
    P(nrbas_max); P(nrbas); P(rwf); P(d2rwf); P(ce); P(ae); P(rmaxwf); P(nbastyp_max); P(nbastyp); P(iwrwf);
    P(iwlbas); P(nbasis); P(norb); P(coef); P(ndet); P(iworbd); P(cdet); P(norda); P(nordb); P(nordc);
    P(nspin2b); P(nocuspb); P(scalek); P(a4); P(b); P(c); P(nloc); P(lpot); P(necp_term); P(necp_power);
    P(ecp_coef); P(ecp_exponent); P(nquad); P(xq0); P(yq0); P(zq0); P(wq);
    printf("sizeof_system %zu\n", sizeof(champ_system_t));
    Q(accepted); Q(proposed); Q(esum); Q(esum2); Q(nflagged);
    printf("sizeof_stats %zu\n", sizeof(champ_sweep_stats));
    V(efin); V(eerr); V(sigma); V(tcorr); V(accept); V(nsamples); V(nflagged);
    printf("sizeof_vmc %zu\n", sizeof(champ_vmc_result));
    D(nwalk); D(pad); D(wsum); D(esum); D(wgsum); D(egsum); D(eest); D(eigv); D(fprod); D(accept); D(nodecr);
    D(nbranch);
    printf("sizeof_dmcstep %zu\n", sizeof(champ_dmc_step_stats));
    R(egave); R(egerr); R(eest); R(accept); R(tau_eff); R(nwalk); R(nwalk_max); R(nbranch); R(nodecr);
    printf("sizeof_dmcres %zu\n", sizeof(champ_dmc_result));
    printf("sr.adiag %zu\nsr.dparm_norm %zu\nsr.cg_iters %zu\nsr.retries %zu\n", offsetof(champ_sr_info, adiag),
           offsetof(champ_sr_info, dparm_norm), offsetof(champ_sr_info, cg_iters), offsetof(champ_sr_info, retries));
    printf("sizeof_sr %zu\n", sizeof(champ_sr_info));
    return 0;
}
