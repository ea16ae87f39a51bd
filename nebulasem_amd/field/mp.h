// The message-passing object the apps construct (src/mp/mp.h MP): on one
// GPU there is one process, so host 0 of 1; printH prefixes the elapsed
// milliseconds and the host as the reference's log lines do
// (mp.cpp:64-80).
#ifndef NSEM_FIELD_MP_H
#define NSEM_FIELD_MP_H

#include "types.h"

class MP {
   public:
    enum { OP_MAX, OP_MIN, OP_SUM, OP_PROD };
    MP(int argc, char* argv[]);
    ~MP();
    static int n_hosts, host_id;
    static bool printOn;
    static bool Terminated;
    static void cleanup();
    static void printH(const char* format, ...);
    static void print(const char* format, ...);
    static bool hasElapsed(const Int delta);
    static void barrier() {}
    // one process: the reduction of one contribution is itself
    template <typename T, typename U = T>
    static void allreduce(T* send, T* recv, int n, Int /*op*/) {
        for (int i = 0; i < n; i++) recv[i] = send[i];
    }
};

#endif
