// Device-memory admission of the serving loop's count jobs.  Host-only and
// free of HIP calls (the caller passes the free HBM it measured), so the
// accounting is tested on any machine (tests/cpp/admit_check.cpp).
//
// A count job holds a ticket.  admit() reserves the job's footprint on the
// ticket; finish() returns exactly that reservation once, whatever way the
// job ended, and adds the bytes of the table the job left resident (0 when it
// failed).  A ticket that was never admitted changes nothing, so a request
// refused before it ran can never release bytes it did not reserve.
#pragma once
#include <cstdint>

struct kc_ticket {
    uint64_t footprint = 0;  // device bytes while the job runs
    bool reserved = false;   // admit() succeeded and finish() has not run
};

struct kc_admission {
    uint64_t budget = 0;    // cap on resident + reserved bytes (0: free HBM alone decides)
    uint64_t margin = 0;    // free HBM kept back for the runtime and for lookups
    uint64_t resident = 0;  // bytes of the tables kept between requests
    uint64_t reserved = 0;  // footprints of admitted jobs still running

    // Whether a job of `bytes` could ever run: it fits the budget with no
    // table resident and the device's total memory.
    bool could_fit(uint64_t bytes, uint64_t total_hbm) const
    {
        return (budget == 0 || bytes <= budget) && bytes + margin <= total_hbm;
    }

    // Admits the ticket's job when its footprint fits the budget next to the
    // resident tables and the running jobs, and the HBM measured free now
    // (free_hbm) next to the running jobs' reservations (some of which may
    // already be allocated and so counted twice: the test errs on the safe
    // side).
    bool admit(kc_ticket &tk, uint64_t free_hbm)
    {
        if (tk.reserved) return true;
        if (budget && resident + reserved + tk.footprint > budget) return false;
        if (reserved + tk.footprint + margin > free_hbm) return false;
        reserved += tk.footprint;
        tk.reserved = true;
        return true;
    }

    // The job ended; table_bytes is what it left resident (0 on failure).
    // Returns whether the ticket held a reservation.
    bool finish(kc_ticket &tk, uint64_t table_bytes)
    {
        if (!tk.reserved) return false;
        reserved -= tk.footprint;
        resident += table_bytes;
        tk.reserved = false;
        return true;
    }

    // A resident table was dropped or evicted.
    void drop(uint64_t table_bytes) { resident -= table_bytes < resident ? table_bytes : resident; }
};
