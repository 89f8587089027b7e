// Force libstdc++'s condition_variable onto pthread_cond_timedwait: GCC 11's ThreadSanitizer
// does not intercept pthread_cond_clockwait, and without the interception it sees a waiting
// thread keep its mutex (false "double lock" and data-race reports).
#include <bits/c++config.h>
#undef _GLIBCXX_USE_PTHREAD_COND_CLOCKWAIT
