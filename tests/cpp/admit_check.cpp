// kc_admission (the serving loop's HBM admission): reservations are
// returned exactly once whatever way a job ends, a ticket that was never
// admitted changes nothing, and resident + reserved bytes return to their
// starting value once every table is dropped -- including the sequence of
// two count jobs on one table id (the second refused without a reservation)
// with the table dropped in between.
#include <cstdint>
#include <cstdio>

#include "kc_admit.h"

static int fail(const char *what)
{
    std::printf("FAIL %s\n", what);
    return 1;
}

int main()
{
    const uint64_t G = 1ull << 30;
    kc_admission a;
    a.budget = 0;
    a.margin = G;
    // admission by measured free HBM
    kc_ticket j1{10 * G, false}, j2{10 * G, false};
    if (!a.admit(j1, 40 * G)) return fail("fits free HBM");
    if (a.admit(j2, 20 * G)) return fail("reservation of the running job counts against free HBM");
    if (a.reserved != 10 * G) return fail("refused job reserves nothing");
    // j1 finishes leaving a 4 GiB table; j2 (same table id) was never admitted
    if (!a.finish(j1, 4 * G)) return fail("admitted job returns its reservation");
    if (a.finish(j1, 4 * G)) return fail("second finish is a no-op");
    if (a.finish(j2, 0)) return fail("unadmitted ticket releases nothing");
    if (a.reserved != 0 || a.resident != 4 * G) return fail("table resident after the job");
    a.drop(4 * G);
    if (a.reserved != 0 || a.resident != 0) return fail("back to the start after the drop");
    // a failed job leaves nothing resident
    kc_ticket j3{5 * G, false};
    if (!a.admit(j3, 100 * G) || !a.finish(j3, 0) || a.resident || a.reserved) return fail("failed job");
    // the budget caps resident + reserved bytes
    a.budget = 12 * G;
    kc_ticket j4{8 * G, false}, j5{8 * G, false};
    if (!a.admit(j4, 100 * G)) return fail("within budget");
    if (a.admit(j5, 100 * G)) return fail("over budget");
    a.finish(j4, 6 * G);
    if (a.admit(j5, 100 * G)) return fail("resident table counts against the budget");
    a.drop(6 * G);
    if (!a.admit(j5, 100 * G)) return fail("fits after the drop");
    a.finish(j5, 0);
    if (a.resident || a.reserved) return fail("end state");
    // jobs that could never run
    if (a.could_fit(13 * G, 288 * G)) return fail("larger than the budget");
    a.budget = 0;
    if (a.could_fit(288 * G, 288 * G)) return fail("larger than the device");
    if (!a.could_fit(200 * G, 288 * G)) return fail("fits the device");
    std::printf("ok\n");
    return 0;
}
