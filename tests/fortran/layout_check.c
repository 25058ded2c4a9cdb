/* offsetof() of the C declarations in include/champ_gpu.h, printed in the
 * same "name offset" form as layout_check.f90. */
#include <stddef.h>
#include <stdio.h>
#include "champ_gpu.h"

#define P(f) printf("%s %zu\n", #f, offsetof(champ_system_t, f))
#define Q(f) printf("stats.%s %zu\n", #f, offsetof(champ_sweep_stats, f))

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
    return 0;
}
