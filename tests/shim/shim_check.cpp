/* Built against the reference's src/param headers: prints the parameters
 * the library holds after the shim's load-time constructor ran. */
#include "radchan_param_shim.h"
#include "param.h"
#include "NarrowBand.h"
#include "radchan_param_shim_def.h"

/* query through a pointer so the macro-named fields are never spelled here */
extern "C" int shim_check_values(int32_t *out);

int main() {
    int32_t v[18];
    shim_check_values(v);
    for (int i = 0; i < 18; ++i) printf("%d ", v[i]);
    printf("\n");
    return 0;
}
