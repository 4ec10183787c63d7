// fdns_host.cpp -- host orchestration and C-ABI of the MI355X time-stepping
// core: scheme registry, per-shape constants, the factor-set cache, run
// lifecycle, batched ensembles and the serving pool.
#include "../../include/fdns.h"
#include "fdns_internal.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

using namespace fdns;

namespace {

thread_local std::string g_err;
thread_local const char* g_err_lit = nullptr;   // a message that needs no copy
std::mutex g_mu;   // guards the shape and factor caches

int fail(int code, const std::string& msg)
{
    g_err_lit = nullptr;
    g_err = msg;
    return code;
}

// For the allocation-free entry point: records a static message, no copy.
int fail_lit(int code, const char* msg)
{
    g_err.clear();
    g_err_lit = msg;
    return code;
}

#define HIPCHK(x)                                                                   \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) return fail(FDNS_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ------------------------------------------------------------------ schemes
struct Scheme {
    int kind = SK_IMEX;
    int imp = IMP_CN;
    int n_imp = 0, n_exp = 0;
    std::string exp;
    int S = 0;
    double aD[MAXS * MAXS] = {0}, aA[MAXS * MAXS] = {0};
    double bD[MAXS] = {0}, bA[MAXS] = {0};
    int diag = 1;          // 0-based diagonal index used for the factors
    int nl_col[MAXS] = {0};
    int fa_col[MAXS] = {0};
    int need_fd[MAXS] = {0};
    int sa = 1;            // stiffly accurate (Assembly_stage_SA) or not (Assembly_stage)
    int diag_diff = 0;     // every stage solved with its own diagonal's operator
};

void set_rows(double* dst, const double* rows, int S)
{
    for (int i = 0; i < S; ++i)
        for (int j = 0; j < S; ++j) dst[i * MAXS + j] = rows[i * S + j];
}

// The stiffly accurate IMEXRK tableaux with diag_diff = .FALSE. and
// diag_index = 2 (rhs_update_wts_imp / rhs_update_wts_exp of timeschemes.F90,
// implicit / explicit case at the lines given): aD and aA row by row; their
// last rows are b.  Every stage k >= 2 is solved with the aD(2,2) operator,
// also for FW53, whose diagonal is not uniform (as the reference runs it).
struct RkTableau {
    const char* name;
    int S;
    std::vector<double> aD, aA;
};

const std::vector<RkTableau>& rk_tableaux()
{
    static const std::vector<RkTableau> T = [] {
        const double g = (2.0 - std::sqrt(2.0)) / 2.0, d = 1.0 - 1.0 / g / 2.0;    // ARS222 gamma, delta
        const double r2 = std::sqrt(2.0), th = 1.0 / 3.0, tq = 3.0 / 4.0;           // CK222
        const double s6 = 1.0 / 6.0;
        return std::vector<RkTableau>{
            {"ARS222", 3,   // :1066 / :1741
             {0, 0, 0, 0, g, 0, 0, 1.0 - g, g},
             {0, 0, 0, g, 0, 0, d, 1.0 - d, 0}},
            {"PC2", 4,      // :1168 / :1866
             {0, 0, 0, 0, 0.5, 0.5, 0, 0, 0.5, 0, 0.5, 0, 0.5, 0, 0, 0.5},
             {0, 0, 0, 0, 1.0, 0, 0, 0, 0.5, 0.5, 0, 0, 0.5, 0, 0.5, 0}},
            {"SSP332", 4,   // :1183 / :1884
             {0, 0, 0, 0, 0, 0.5, 0, 0, 0, 0.5, 0.5, 0, 0, 1.0, -0.5, 0.5},
             {0, 0, 0, 0, 0.5, 0, 0, 0, 0.5, 0.5, 0, 0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 0}},
            {"SSP543", 6,   // :1198 / :1902
             {0, 0, 0, 0, 0, 0,
              0, 0.5, 0, 0, 0, 0,
              0, 0.5, 0.5, 0, 0, 0,
              0, 0.5, -0.5, 0.5, 0, 0,
              0, 2.0, 0.5, -3.0, 0.5, 0,
              0, 2.0 / 3.0, -1.0 / 3.0, 0, s6, 0.5},
             {0, 0, 0, 0, 0, 0,
              0.5, 0, 0, 0, 0, 0,
              0.5, 0.5, 0, 0, 0, 0,
              s6, s6, s6, 0, 0, 0,
              0, 0, 0, 0, 0, 0,
              s6, s6, s6, 0.5, 0, 0}},
            {"ARS443", 5,   // :1215 / :1919
             {0, 0, 0, 0, 0,
              0, 0.5, 0, 0, 0,
              0, s6, 0.5, 0, 0,
              0, -0.5, 0.5, 0.5, 0,
              0, 1.5, -1.5, 0.5, 0.5},
             {0, 0, 0, 0, 0,
              0.5, 0, 0, 0, 0,
              11.0 / 18.0, 1.0 / 18.0, 0, 0, 0,
              5.0 / 6.0, -5.0 / 6.0, 0.5, 0, 0,
              0.25, 7.0 / 4.0, 0.75, -7.0 / 4.0, 0}},
            {"BPR353", 5,   // :1277 / :1995
             {0, 0, 0, 0, 0,
              0.5, 0.5, 0, 0, 0,
              5.0 / 18.0, -1.0 / 9.0, 0.5, 0, 0,
              0.5, 0, 0, 0.5, 0,
              0.25, 0, 0.75, -0.5, 0.5},
             {0, 0, 0, 0, 0,
              1.0, 0, 0, 0, 0,
              4.0 / 9.0, 2.0 / 9.0, 0, 0, 0,
              0.25, 0, 0.75, 0, 0,
              0.25, 0, 0.75, 0, 0}},
            {"FW53", 5,     // :1410 / :2151
             {0, 0, 0, 0, 0,
              0, 1.0, 0, 0, 0,
              0.5, 0, 0.5, 0, 0,
              0.25, -0.25, 0, 0.5, 0,
              s6, s6, -s6, 4.0 / 6.0, s6},
             {0, 0, 0, 0, 0,
              1.0, 0, 0, 0, 0,
              0.5, 0.5, 0, 0, 0,
              0.25, 0.25, 0, 0, 0,
              s6, s6, 0, 4.0 / 6.0, 0}},
            {"CK222", 3,    // :1430 / :2173
             {0, 0, 0,
              -th + r2 * 0.5, 1.0 - r2 * 0.5, 0,
              tq - r2 * 0.25, -tq + r2 * 0.75, 1.0 - r2 * 0.5},
             {0, 0, 0, 2.0 / 3.0, 0, 0, 0.25, 0.75, 0}},
            {"BPR442", 5,   // :1449 / :2194
             {0, 0, 0, 0, 0,
              0, 0.25, 0, 0, 0,
              0, 0, 0.25, 0, 0,
              0, 1.0 / 24.0, 11.0 / 24.0, 0.25, 0,
              0, 11.0 / 24.0, s6, 1.0 / 8.0, 0.25},
             {0, 0, 0, 0, 0,
              0.25, 0, 0, 0, 0,
              13.0 / 4.0, -3.0, 0, 0, 0,
              0.25, 0, 0.5, 0, 0,
              0, 1.0 / 3.0, s6, 0.5, 0}},
        };
    }();
    return T;
}

// The other IMEXRK tableaux, as doubles (row-major aD, bD, aA, bA): the
// reference's constant expressions evaluated in double arithmetic, to 17
// digits.  ckA / ckD are the ars_eqn_check flags the reference ends up with
// (set literally, or by its any-element loop): both set, the step ends with
// Assembly_stage_SA, otherwise with Assembly_stage.  diag_diff / diag_index
// as the reference sets them.
struct RkNumeric {
    const char* name;
    int S, ckA, ckD, diag_diff, diag_index;
    double aD[MAXS * MAXS], bD[MAXS], aA[MAXS * MAXS], bA[MAXS];
};

const RkNumeric RK_NUMERIC[] = {
    // EUL111: rhs_update_wts_imp :1055, rhs_update_wts_exp :1727, init_time_schemes :190
    {"EUL111", 2, 0, 0, 0, 2,
     {0, 0,
      0, 1},
     {0, 1},
     {0, 0,
      1, 0},
     {1, 0}
    },
    // ARS232: rhs_update_wts_imp :1080, rhs_update_wts_exp :1758, init_time_schemes :236
    {"ARS232", 3, 0, 1, 0, 2,
     {0, 0, 0,
      0, 0.29289321881345243, 0,
      0, 0.70710678118654757, 0.29289321881345243},
     {0, 0.70710678118654757, 0.29289321881345243},
     {0, 0, 0,
      0.29289321881345243, 0, 0,
      -0.94280904158206347, 1.9428090415820636, 0},
     {0, 0.70710678118654757, 0.29289321881345243}
    },
    // ARS233: rhs_update_wts_imp :1094, rhs_update_wts_exp :1775, init_time_schemes :260
    {"ARS233", 3, 0, 0, 0, 2,
     {0, 0, 0,
      0, 0.78867513459481275, 0,
      0, -0.57735026918962551, 0.78867513459481275},
     {0, 0.5, 0.5},
     {0, 0, 0,
      0.78867513459481275, 0, 0,
      -0.21132486540518725, 0.42264973081037449, 0},
     {0, 0.5, 0.5}
    },
    // ARS343: rhs_update_wts_imp :1107, rhs_update_wts_exp :1792, init_time_schemes :283
    {"ARS343", 4, 0, 1, 0, 2,
     {0, 0, 0, 0,
      0, 0.435866521508459, 0, 0,
      0, 0.28206673924577053, 0.435866521508459, 0,
      0, 1.2084966491760101, -0.64436317068446924, 0.435866521508459},
     {0, 1.2084966491760101, -0.64436317068446924, 0.435866521508459},
     {0, 0, 0, 0,
      0.435866521508459, 0, 0, 0,
      0.32127888602862781, 0.39665437472560172, 0, 0,
      -0.1058582960718797, 0.55292914803593984, 0.55292914803593984, 0},
     {0, 1.2084966491760101, -0.64436317068446924, 0.435866521508459}
    },
    // MARS343: rhs_update_wts_imp :1139, rhs_update_wts_exp :1830, init_time_schemes :341
    {"MARS343", 4, 0, 1, 0, 2,
     {0, 0, 0, 0,
      0, 0.435866521508458, 0, 0,
      0, 0.28206673924577003, 0.435866521508458, 0,
      0, 1.2084966491760101, -0.64436317068446003, 0.435866521508458},
     {0, 1.2084966491760101, -0.64436317068446003, 0.435866521508458},
     {0, 0, 0, 0,
      0.435866521508458, 0, 0, 0,
      0.53539654030735395, 0.18253672044687499, 0, 0,
      0.63041255815287001, -0.83193390106307996, 1.2015213429102101, 0},
     {0, 1.2084966491760101, -0.64436317068446003, 0.435866521508458}
    },
    // IJ3: rhs_update_wts_imp :1154, rhs_update_wts_exp :1848, init_time_schemes :363
    {"IJ3", 4, 0, 0, 0, 2,
     {0, 0, 0, 0,
      0, 0.78868665109985225, 0, 0,
      0, -0.45533028521586527, 0.78866361854919842, 0,
      0, 1.3659908556475957, -0.36599085564759576, 0},
     {0, 0, 0.75, 0.25},
     {0, 0, 0, 0,
      0, 0, 0, 0,
      0, 0.33333333333333331, 0, 0,
      0, -1, 2, 0},
     {0, 0, 0.75, 0.25}
    },
    // PR433: rhs_update_wts_imp :1235, rhs_update_wts_exp :1946, init_time_schemes :474
    {"PR433", 5, 0, 0, 0, 2,
     {0, 0, 0, 0, 0,
      0, 0.24169426078820999, 0, 0, 0,
      0, -0.24169426078820999, 0.24169426078820999, 0, 0,
      0, 0, 0.75830573921179001, 0.24169426078820999, 0,
      0, 0.06042356519705, 0.1291528696059, 0.068729304408840008, 0.24169426078820999},
     {0, 0, 0.16666666666666666, 0.16666666666666666, 0.66666666666666663},
     {0, 0, 0, 0, 0,
      0, 0, 0, 0, 0,
      0, 0, 0, 0, 0,
      0, 0, 1, 0, 0,
      0, 0, 0.25, 0.25, 0},
     {0, 0, 0.16666666666666666, 0.16666666666666666, 0.66666666666666663}
    },
    // DBM453: rhs_update_wts_imp :1258, rhs_update_wts_exp :1978, init_time_schemes :499
    {"DBM453", 5, 0, 1, 0, 2,
     {0, 0, 0, 0, 0,
      -0.2228498531852541, 0.32591194130117246, 0, 0, 0,
      -0.46801347074080546, 0.8634928422571696, 0.32591194130117246, 0, 0,
      -0.46509906651927418, 0.81063103116959556, 0.6103672675683236, 0.32591194130117246, 0,
      0.87795339639076675, -0.72692641526151547, 0.75204137157372719, -0.22898029400415087, 0.32591194130117246},
     {0.87795339639076675, -0.72692641526151547, 0.75204137157372719, -0.2289802940041509, 0.32591194130117246},
     {0, 0, 0, 0, 0,
      0.10306208811591838, 0, 0, 0, 0,
      -0.94124866143519892, 1.6626399742527356, 0, 0, 0,
      -1.3670975201437765, 1.3815852911016873, 1.2673234025619065, 0, 0,
      -0.81287582068772446, 0.8122373906050574, 0.90644429603699306, 0.094194134045674116, 0},
     {0.87795339639076675, -0.72692641526151547, 0.75204137157372719, -0.2289802940041509, 0.32591194130117246}
    },
    // KC343: rhs_update_wts_imp :1297, rhs_update_wts_exp :2022, init_time_schemes :544
    {"KC343", 4, 0, 1, 0, 2,
     {0, 0, 0, 0,
      0.435866521508459, 0.435866521508459, 0, 0,
      0.25764824606642722, -0.093514767574886248, 0.435866521508459, 0,
      0.18764102434672383, -0.59529747357695495, 0.97178992772177208, 0.435866521508459},
     {0.18764102434672383, -0.59529747357695495, 0.97178992772177208, 0.435866521508459},
     {0, 0, 0, 0,
      0.87173304301691801, 0, 0, 0,
      0.52758901197630037, 0.072410988023699593, 0, 0,
      0.39909600767607012, -0.43755765461351942, 1.0384616469374492, 0},
     {0.18764102434672383, -0.59529747357695495, 0.97178992772177208, 0.435866521508459}
    },
    // KC564: rhs_update_wts_imp :1323, rhs_update_wts_exp :2052, init_time_schemes :566
    {"KC564", 6, 0, 1, 0, 2,
     {0, 0, 0, 0, 0, 0,
      0.25, 0.25, 0, 0, 0, 0,
      0.13777600000000001, -0.055775999999999999, 0.25, 0, 0, 0,
      0.14463686602698217, -0.22393190761334475, 0.44929504158636258, 0.25, 0, 0,
      0.098258783283564771, -0.59154424281967044, 0.81012105382829958, 0.28316440570780599, 0.25, 0,
      0.15791629516167136, 0, 0.18675894052400077, 0.68056529530933463, -0.27524053099500667, 0.25},
     {0.15791629516167136, 0, 0.18675894052400077, 0.68056529530933463, -0.27524053099500667, 0.25},
     {0, 0, 0, 0, 0, 0,
      0.5, 0, 0, 0, 0, 0,
      0.221776, 0.110224, 0, 0, 0, 0,
      -0.04884659515311858, -0.177720652326401, 0.84656724747951961, 0, 0, 0,
      -0.15541685842491548, -0.3567050098221991, 1.0587258798684427, 0.30339598837867193, 0, 0,
      0.20142435067267633, 0.0087420578429041849, 0.15993995707168115, 0.40382906052207751, 0.22606457389066084, 0},
     {0.15791629516167136, 0, 0.18675894052400077, 0.68056529530933463, -0.27524053099500667, 0.25}
    },
    // KC774: rhs_update_wts_imp :1342, rhs_update_wts_exp :2076, init_time_schemes :588
    {"KC774", 7, 0, 1, 0, 2,
     {0, 0, 0, 0, 0, 0, 0,
      0.1235, 0.1235, 0, 0, 0, 0, 0,
      0.14907768747653863, 0.14907768747653863, 0.1235, 0, 0, 0, 0,
      0.12483442871739439, 0.12483442871739439, -0.038168857434788782, 0.1235, 0, 0, 0,
      -0.073031940302180909, -0.073031940302180909, -0.24343568716014671, 0.34099956776450852, 0.1235, 0, 0,
      -0.15296500088128806, -0.15296500088128806, 0.072205620474335874, 0.40430630248551713, 0.40591807880272318, 0.1235, 0,
      0, 0, 0.51611072831742366, -0.14606356393857081, 0.23473048589019332, 0.27172234973095377, 0.1235},
     {0, 0, 0.51611072831742366, -0.14606356393857081, 0.23473048589019332, 0.27172234973095377, 0.1235},
     {0, 0, 0, 0, 0, 0, 0,
      0.247, 0, 0, 0, 0, 0, 0,
      0.061749999999999999, 0.35990537495307723, 0, 0, 0, 0, 0,
      0.05301658458687121, 0.35949264529328429, -0.077509229880155461, 0, 0, 0, 0,
      0.058417159447800002, -0.16313824817772324, -0.19732090979798411, 0.37704199852790737, 0, 0, 0,
      0.53853032270810797, -0.45497746895916669, 1.2562905623429941, -0.47828452721130055, -0.16155888888063494, 0, 0,
      0.23221715782277083, 0.23221715782277083, -6.809994375038098, 7.3618585524244216, -1.3748790779406981, 1.3585805849088326, 0},
     {0, 0, 0.51611072831742366, -0.14606356393857081, 0.23473048589019332, 0.27172234973095377, 0.1235}
    },
    // KC885: rhs_update_wts_imp :1373, rhs_update_wts_exp :2110, init_time_schemes :610
    {"KC885", 8, 0, 1, 0, 2,
     {0, 0, 0, 0, 0, 0, 0, 0,
      0.22222222222222221, 0.22222222222222221, 0, 0, 0, 0, 0, 0,
      0.26824595137478835, 0.26824595137478835, 0.22222222222222221, 0, 0, 0, 0, 0,
      -0.057945592237231995, -0.057945592237231995, 0.0089383968162837328, 0.22222222222222221, 0, 0, 0, 0,
      -0.043305287723547685, -0.043305287723547685, -0.034013891077568637, 0.25515937270676026, 0.22222222222222221, 0, 0, 0,
      0.13179599023759678, 0.13179599023759678, -0.032376726277862332, 0.12385474427672251, 0.14270777930372408, 0.22222222222222221, 0, 0,
      0.30932282100434261, 0.30932282100434261, -0.68291992723367922, -0.058822756149695461, -0.041308613833499437, 0.89718343298596659, 0.22222222222222221, 0,
      0, 0, 0.17366253573581261, 0.25479166260812353, 0.24190176845094791, 0.30740485830222825, -0.19998304731933453, 0.22222222222222221},
     {0, 0, 0.17366253573581261, 0.25479166260812353, 0.24190176845094791, 0.30740485830222825, -0.19998304731933453, 0.22222222222222221},
     {0, 0, 0, 0, 0, 0, 0, 0,
      0.44444444444444442, 0, 0, 0, 0, 0, 0, 0,
      0.1111111111111111, 0.64760301386068775, 0, 0, 0, 0, 0, 0,
      0.091829866647747907, 0.035448567517799241, -0.012008999601505184, 0, 0, 0, 0, 0,
      -0.34252354516023137, -0.26767785943050182, 0.11056894178117282, 0.85638959121387881, 0, 0, 0, 0,
      -0.0097722828790043955, 0.21070865398661751, 0.075924120912175361, 0.20765518596381696, 0.23548432201639455, 0, 0, 0,
      0.46686370681500694, 1.2903598800650855, 0.37840596884419414, -0.56345584032826157, -0.2883238346202236, -0.32884988077580141, 0, 0,
      0.61439671625166914, 0.61439671625166914, 0.31747780106686158, -0.71215206239529361, 0.11498708015310211, 0.09139031575415682, -0.040496567082165244, 0},
     {0, 0, 0.17366253573581261, 0.25479166260812353, 0.24190176845094791, 0.30740485830222825, -0.19998304731933453, 0.22222222222222221}
    },
    // CFN343: rhs_update_wts_imp :1465, rhs_update_wts_exp :2217, init_time_schemes :698
    {"CFN343", 4, 0, 1, 0, 2,
     {0, 0, 0, 0,
      0, 0.435866521508459, 0, 0,
      0, 0.28206673924577053, 0.435866521508459, 0,
      0, 1.2084966491760101, -0.64436317068446902, 0.435866521508459},
     {0, 1.2084966491760101, -0.64436317068446902, 0.435866521508459},
     {0, 0, 0, 0,
      0.435866521508459, 0, 0, 0,
      1.0679332607542293, -0.34999999999999998, 0, 0,
      0, 1.989175724679846, -0.98917572467984605, 0},
     {0, 1.2084966491760101, -0.64436317068446902, 0.435866521508459}
    },
    // BHR553: rhs_update_wts_imp :1480, rhs_update_wts_exp :2235, init_time_schemes :721
    {"BHR553", 5, 0, 1, 0, 2,
     {0, 0, 0, 0, 0,
      0.43586652150848199, 0.43586652150848199, 0, 0, 0,
      0.43586652150848199, -6.9635084214665835e-14, 0.43586652150848199, 0, 0,
      -0.066758687019583762, -6.1296596594174833e-13, 1.9711047406203948, 0.43586652150848199, 0,
      0.41289804281247428, 0, 0.19733989037886918, -0.04610445469982568, 0.43586652150848199},
     {0.41289804281247428, 0, 0.19733989037886918, -0.04610445469982568, 0.43586652150848199},
     {0, 0, 0, 0, 0,
      0.87173304301696397, 0, 0, 0, 0,
      0.43586652150848199, 0.43586652150848199, 0, 0, 0,
      -0.80099845306562878, 0, 3.1412110281743093, 0, 0,
      0.35675320777963954, -0.19733989037886918, 0.88194884139379082, -0.041362158794624231, 0},
     {0.41289804281247428, 0, 0.19733989037886918, -0.04610445469982568, 0.43586652150848199}
    },
    // CFN564: rhs_update_wts_imp :1575, rhs_update_wts_exp :2349, init_time_schemes :918
    {"CFN564", 6, 0, 1, 0, 2,
     {0, 0, 0, 0, 0, 0,
      0, 0.25, 0, 0, 0, 0,
      0, 0.5, 0.25, 0, 0, 0,
      0, 0.34000000000000002, -0.040000000000000001, 0.25, 0, 0,
      0, 0.2727941176470588, -0.050367647058823531, 0.027573529411764705, 0.25, 0,
      0, 1.0416666666666667, -1.0208333333333333, 7.8125, -7.083333333333333, 0.25},
     {0, 1.0416666666666667, -1.0208333333333333, 7.8125, -7.083333333333333, 0.25},
     {0, 0, 0, 0, 0, 0,
      0.25, 0, 0, 0, 0, 0,
      -0.25, 1, 0, 0, 0, 0,
      -0.13, 0.57333333333333336, 0.10666666666666667, 0, 0, 0,
      -0.070588235294117646, 0.49411764705882355, 0.13161764705882353, -0.05514705882352941, 0, 0,
      0, 3.2916666666666665, -0.625, 12.5, -14.166666666666666, 0},
     {0, 1.0416666666666667, -1.0208333333333333, 7.8125, -7.083333333333333, 0.25}
    },
    // stages solved with their own diagonal (diag_diff = .TRUE.,
    // fourierloop_imexrk.F90:134-158)
    // SMR333: rhs_update_wts_imp :1122, rhs_update_wts_exp :1810, init_time_schemes :308
    {"SMR333", 4, 1, 1, 1, 2,
     {0, 0, 0, 0,
      0.30208333333333331, 0.23125000000000001, 0, 0,
      0.30208333333333331, 0.15625, 0.20833333333333334, 0,
      0.30208333333333331, 0.15625, 0.375, 0.16666666666666666},
     {0.30208333333333331, 0.15625, 0.375, 0.16666666666666666},
     {0, 0, 0, 0,
      0.53333333333333333, 0, 0, 0,
      0.25, 0.41666666666666669, 0, 0,
      0.25, 0, 0.75, 0},
     {0.25, 0, 0.75, 0}
    },
    // CB3e: rhs_update_wts_imp :1496, rhs_update_wts_exp :2254, init_time_schemes :765
    {"CB3e", 4, 0, 1, 1, 3,
     {0, 0, 0, 0,
      0, 0.33333333333333331, 0, 0,
      0, 0.5, 0.5, 0,
      0, 0.75, -0.25, 0.5},
     {0, 0.75, -0.25, 0.5},
     {0, 0, 0, 0,
      0.33333333333333331, 0, 0, 0,
      0, 1, 0, 0,
      0, 0.75, 0.25, 0},
     {0, 0.75, -0.25, 0.5}
    },
    // CB3f: rhs_update_wts_imp :1511, rhs_update_wts_exp :2272, init_time_schemes :787
    {"CB3f", 4, 0, 1, 1, 2,
     {0, 0, 0, 0,
      0.48999999999999999, 0.48999999999999999, 0, 0,
      -0.71803538544242407, -0.031405761082656344, 0.78944114652508046, 0,
      -3.6143737489279602, 0.11126152710363123, 4.2834880399852997, 0.21962418183902943},
     {-3.6143737489279602, 0.11126152710363123, 4.2834880399852997, 0.21962418183902943},
     {0, 0, 0, 0,
      0.97999999999999998, 0, 0, 0,
      0.020449687951736302, 0.019550312048263699, 0, 0,
      -3.6143737489279602, 0.21342496834978161, 4.4009487805781786, 0},
     {-3.6143737489279602, 0.11126152710363123, 4.2834880399852997, 0.21962418183902943}
    },
    // CB4: rhs_update_wts_imp :1526, rhs_update_wts_exp :2290, init_time_schemes :824
    {"CB4", 6, 0, 1, 1, 2,
     {0, 0, 0, 0, 0, 0,
      0.125, 0.125, 0, 0, 0, 0,
      0.22486299228804568, 0.22520551156793148, 0.29993149614402281, 0, 0, 0,
      0.1685018686835717, -0.29862732345987331, -0.1185405724697194, 0.62366602724602105, 0, 0,
      0.1685018686835717, 0.14353694052880869, -0.11698581780682539, -0.19505299140555502, 0.5, 0,
      0.1685018686835717, 0.14353694052880869, -0.15821855666404905, -0.48280879550857242, 1.1303111245573836, 0.19867741840285749},
     {0.1685018686835717, 0.14353694052880869, -0.15821855666404905, -0.48280879550857242, 1.1303111245573836, 0.19867741840285749},
     {0, 0, 0, 0, 0, 0,
      0.25, 0, 0, 0, 0, 0,
      0.15321917472916377, 0.5967808252708362, 0, 0, 0, 0,
      0.1685018686835717, 0.12100822974370064, 0.08548990157272765, 0, 0, 0,
      0.1685018686835717, 0.14353694052880869, 0.075161004492142164, 0.11280018629547743, 0, 0,
      0.1685018686835717, 0.14353694052880869, -0.15821855666404905, -0.64963614327365304, 1.4958158907253216, 0},
     {0.1685018686835717, 0.14353694052880869, -0.15821855666404905, -0.48280879550857242, 1.1303111245573836, 0.19867741840285749}
    },
    // LZ3L1: rhs_update_wts_imp :1543, rhs_update_wts_exp :2309, init_time_schemes :873
    {"LZ3L1", 5, 1, 1, 1, 2,
     {0, 0, 0, 0, 0,
      0.14999999999999999, 0.10000000000000001, 0, 0, 0,
      0.90000000000000002, -1.3, 0.90000000000000002, 0, 0,
      1.7, -2.75, 1.5, 0.29999999999999999, 0,
      1, -3.3333333333333335, 5.666666666666667, -3.3333333333333335, 1},
     {1, -3.3333333333333335, 5.666666666666667, -3.3333333333333335, 1},
     {0, 0, 0, 0, 0,
      0.25, 0, 0, 0, 0,
      0, 0.5, 0, 0, 0,
      -0.5, 1.25, 0, 0, 0,
      0, 0.66666666666666663, -0.33333333333333331, 0.66666666666666663, 0},
     {0, 0.66666666666666663, -0.33333333333333331, 0.66666666666666663, 0}
    },
    // LZ3A4a: rhs_update_wts_imp :1559, rhs_update_wts_exp :2328, init_time_schemes :895
    {"LZ3A4a", 5, 1, 1, 1, 3,
     {0, 0, 0, 0, 0,
      0.5, 0, 0, 0, 0,
      0.25, -0.75, 1, 0, 0,
      0, -3, 4, 0, 0,
      0.16666666666666666, 0, 0.66666666666666663, -0.5, 0.66666666666666663},
     {0.16666666666666666, 0, 0.66666666666666663, -0.5, 0.66666666666666663},
     {0, 0, 0, 0, 0,
      0.5, 0, 0, 0, 0,
      0.25, 0.25, 0, 0, 0,
      0, 1, 0, 0, 0,
      0.16666666666666666, 0, 0.66666666666666663, 0.16666666666666666, 0},
     {0.16666666666666666, 0, 0.66666666666666663, 0.16666666666666666, 0}
    },
    // LZ4A1: rhs_update_wts_imp :1592, rhs_update_wts_exp :2369, init_time_schemes :940
    {"LZ4A1", 7, 1, 1, 1, 2,
     {0, 0, 0, 0, 0, 0, 0,
      -0.16666666666666666, 0.5, 0, 0, 0, 0, 0,
      0.16666666666666666, -0.33333333333333331, 0.5, 0, 0, 0, 0,
      0.375, -0.375, 0, 0.5, 0, 0, 0,
      0.125, 0, 0.375, -0.5, 0.5, 0, 0,
      -0.5, 0, 3, -2, 0, 0.5, 0,
      0.16666666666666666, 0, 0, 0, 0.66666666666666663, -0.5, 0.66666666666666663},
     {0.16666666666666666, 0, 0, 0, 0.66666666666666663, -0.5, 0.66666666666666663},
     {0, 0, 0, 0, 0, 0, 0,
      0.33333333333333331, 0, 0, 0, 0, 0, 0,
      0.16666666666666666, 0.16666666666666666, 0, 0, 0, 0, 0,
      0.125, 0, 0.375, 0, 0, 0, 0,
      0.125, 0, 0.375, 0, 0, 0, 0,
      0.5, 0, -1.5, 1, 1, 0, 0,
      0.16666666666666666, 0, 0, 0, 0.66666666666666663, 0.16666666666666666, 0},
     {0.16666666666666666, 0, 0, 0, 0.66666666666666663, 0.16666666666666666, 0}
    },
};

// Registry of timeschemes.F90 (init_time_schemes :42-973) and the IMEXRK
// tableaux above.
int scheme_lookup(const char* type, const char* imp, const char* exp, Scheme& sc)
{
    sc = Scheme();
    sc.exp = exp;
    const std::string t(type), i(imp), e(exp);
    if (t == "IMEX") {
        sc.kind = SK_IMEX;
        if (i == "CN") {
            sc.imp = IMP_CN; sc.n_imp = 2;
            if (e == "AB2") sc.n_exp = 2;
            else if (e == "AB3") sc.n_exp = 3;
            else return FDNS_ESCHEME;
        } else if (i == "BDF2" && e == "AB2") { sc.imp = IMP_BDF2; sc.n_imp = 2; sc.n_exp = 2; }
        else if (i == "BDF3" && e == "AB3") { sc.imp = IMP_BDF3; sc.n_imp = 3; sc.n_exp = 3; }
        else if (i == "BDF4" && e == "AB4") { sc.imp = IMP_BDF4; sc.n_imp = 4; sc.n_exp = 4; }
        else return FDNS_ESCHEME;
        return 0;
    }
    if (t == "IMEXRK") {
        sc.kind = SK_IMEXRK;
        const RkTableau* tab = nullptr;
        const RkNumeric* num = nullptr;
        if (e.compare(0, 3, "exp") == 0 && e.substr(3) == i) {
            for (const auto& k : rk_tableaux())
                if (i == k.name) tab = &k;
            for (const auto& k : RK_NUMERIC)
                if (i == k.name) num = &k;
        }
        bool ckA = false, ckD = false;
        if (tab) {
            sc.S = tab->S;
            set_rows(sc.aD, tab->aD.data(), sc.S);
            set_rows(sc.aA, tab->aA.data(), sc.S);
            for (int k = 0; k < sc.S; ++k) {   // b: the last rows (stiffly accurate)
                sc.bD[k] = tab->aD[(size_t)(sc.S - 1) * sc.S + k];
                sc.bA[k] = tab->aA[(size_t)(sc.S - 1) * sc.S + k];
            }
            // ars_eqn_check_A/D (timeschemes.F90:1226-1230): any element of
            // the last row equal to b
            for (int k = 0; k < sc.S; ++k) {
                if (sc.aD[(sc.S - 1) * MAXS + k] == sc.bD[k]) ckD = true;
                if (sc.aA[(sc.S - 1) * MAXS + k] == sc.bA[k]) ckA = true;
            }
        } else if (num) {
            sc.S = num->S;
            set_rows(sc.aD, num->aD, sc.S);
            set_rows(sc.aA, num->aA, sc.S);
            std::memcpy(sc.bD, num->bD, sizeof(double) * (size_t)sc.S);
            std::memcpy(sc.bA, num->bA, sizeof(double) * (size_t)sc.S);
            ckA = num->ckA != 0;
            ckD = num->ckD != 0;
            sc.diag_diff = num->diag_diff;
        } else {
            return FDNS_ESCHEME;   // diag_diff schemes and the RK family are not built
        }
        sc.sa = (ckA && ckD) ? 1 : 0;
        sc.n_imp = sc.S; sc.n_exp = sc.S;
        sc.diag = num ? num->diag_index - 1 : 1;   // diag_index (1-based) - 1
        for (int k = 0; k < sc.S; ++k) {
            double s = 0.0;
            for (int r = 0; r < sc.S; ++r) s += sc.aA[r * MAXS + k];
            sc.nl_col[k] = (s + sc.bA[k] != 0.0) ? 1 : 0;
            // stage k's explicit terms are read later: a nonzero aA(:,k) or
            // bA(k).  Where the column sums to zero without being zero
            // (CB3e's third column), the pass is skipped and the terms come
            // from the products of the last pass, as RHS_construct_stage
            // builds them from the products it is given
            // (steptime_imexrk.F90:153-166)
            int fa = sc.bA[k] != 0.0 ? 1 : 0;
            for (int r = k + 1; r < sc.S; ++r) fa |= (sc.aA[r * MAXS + k] != 0.0);
            sc.fa_col[k] = fa;
            // stage k's diffusion terms are read by a later stage, or by
            // Assembly_stage through bD
            int need = (!sc.sa && sc.bD[k] != 0.0) ? 1 : 0;
            for (int r = k + 1; r < sc.S; ++r) need |= (sc.aD[r * MAXS + k] != 0.0);
            sc.need_fd[k] = need;
        }
        return 0;
    }
    if (t == "RK") {
        // the explicit path (time_scheme_imp = 'NONE', timeschemes.F90:158-186,
        // 1701-1721): butcher_a / butcher_b in aA / bA.  RK4's a(2,3) = 1/2
        // above the diagonal is the reference's; the stage sums stop below
        // the diagonal, so it is never read.  n_order_tscheme_imp =
        // n_order_tscheme_exp = the stage count (the checkpoint header).
        if (i != "NONE") return FDNS_ESCHEME;
        sc.kind = SK_RK;
        if (e == "RK2") {
            sc.S = 2;
            sc.aA[1 * MAXS + 0] = 0.5;
            sc.bA[0] = 0.0; sc.bA[1] = 1.0;
        } else if (e == "RK4") {
            sc.S = 4;
            sc.aA[1 * MAXS + 0] = 0.5; sc.aA[1 * MAXS + 2] = 0.5;
            sc.aA[2 * MAXS + 1] = 0.5;
            sc.aA[3 * MAXS + 2] = 1.0;
            sc.bA[0] = 1.0 / 6.0; sc.bA[1] = 1.0 / 3.0; sc.bA[2] = 1.0 / 3.0; sc.bA[3] = 1.0 / 6.0;
        } else {
            return FDNS_ESCHEME;
        }
        sc.n_imp = sc.S; sc.n_exp = sc.S;
        sc.sa = 0;
        return 0;
    }
    return FDNS_ESCHEME;
}

// rhs_update_wts_imp / rhs_update_wts_exp (timeschemes.F90:993-1052,
// 1614-1695): the variable-step formulas evaluated with dt_array (dta[0] the
// newest dt).
void imex_weights(const Scheme& sc, const double* dta, double& wl, double* wi, double* we)
{
    for (int k = 0; k < MAXH; ++k) { wi[k] = 0.0; we[k] = 0.0; }
    if (sc.imp == IMP_CN) {
        wi[0] = 0.5;
        wl = 0.5;
        if (sc.exp == "AB2") {
            we[0] = 1.0 + 0.5 * dta[0] / dta[1];
            we[1] = -0.5 * dta[0] / dta[1];
        } else {
            we[1] = (dta[0] * dta[2] * 0.5 + (1.0 / 6.0) * dta[0] * dta[0]) / (dta[1] * dta[1] * 0.5 - dta[1] * dta[2]);
            we[2] = (-dta[0] * 0.5 - dta[1] * we[1]) * 0.5 / (dta[2]);
            we[0] = 1.0 - we[1] - we[2];
        }
    } else if (sc.imp == IMP_BDF2) {
        const double r1 = dta[1] / dta[0];
        const double a0 = (2.0 + r1) / (1.0 + r1);
        const double a1 = 1.0 + 1.0 / r1;
        const double a2 = -1.0 * (1.0 / (r1) / (1.0 + r1));
        wl = 1.0 / a0; wi[0] = a1 / a0; wi[1] = a2 / a0;
        const double b0 = 1.0 + 1.0 / r1, b1 = -1.0 / r1;
        we[0] = b0 / a0; we[1] = b1 / a0;
    } else if (sc.imp == IMP_BDF3) {
        const double r1 = dta[1] / dta[0];
        const double r2 = dta[2] / dta[0];
        const double a0 = (1.0 + 1.0 / (1.0 + r1) + 1.0 / (1.0 + r1 + r2));
        const double a1 = (((1.0 + r1) * (1.0 + r1 + r2)) / (r1 * (r1 + r2)));
        const double a2 = (-1.0 * (1.0 + r1 + r2) / (r1 * r2 * (1.0 + r1)));
        const double a3 = ((1.0 + r1) / (r2 * (r1 + r2) * (1.0 + r1 + r2)));
        wl = 1.0 / a0; wi[0] = a1 / a0; wi[1] = a2 / a0; wi[2] = a3 / a0;
        const double b0 = 1.0 * ((1.0 + r1) * (1.0 + r1 + r2)) / (r1 * (r1 + r2));
        const double b1 = 1.0 * (-(1.0 + r1 + r2) / (r1 * r2));
        const double b2 = 1.0 * (1.0 + r1) / (r2 * (r1 + r2));
        we[0] = b0 / a0; we[1] = b1 / a0; we[2] = b2 / a0;
    } else {
        const double one = 1.0;
        const double r1 = dta[0] / dta[1], r2 = dta[1] / dta[2], r3 = dta[2] / dta[3];
        const double c1 = one + r3 * (one + r2);
        const double c2 = one + r2 * (one + r1);
        const double c3 = one + r3 * c2;
        const double a0 = one + r1 / (one + r1) + r2 * r1 / c2 + r3 * r2 * r1 / c3;
        const double a1 = -one - r1 * (one + r2 * (one + r1) * (one + r3 * c2 / c1) / (one + r2));
        const double a2 = r1 * (r1 / (one + r1) + r2 * r1 * (c3 + r3) / (one + r3));
        const double a3 = -std::pow(r2, 3) * std::pow(r1, 2) * (one + r1) / (one + r2) * c3 / c2;
        const double a4 = (one + r1) / (one + r3) * c2 / c1 / c3 * (std::pow(r3, 4) * std::pow(r2, 3) * std::pow(r1, 2));
        wl = 1.0 / a0;
        wi[0] = -a1 / a0; wi[1] = -a2 / a0; wi[2] = -a3 / a0; wi[3] = -a4 / a0;
        const double b3 = -std::pow(r3, 3) * std::pow(r2, 2) * r1 * (one + r1) / (one + r3) * c2 / c1;
        const double b2 = std::pow(r2, 2) * r1 * (one + r1) / (one + r2) * c3;
        const double b1 = -c2 * c3 * r1 / (one + r3);
        const double b0 = r2 * (one + r1) / (one + r2) * ((one + r1) * (c3 + r3) + (one + r3) / r2) / c1;
        we[0] = b0 / a0; we[1] = b1 / a0; we[2] = b2 / a0; we[3] = b3 / a0;
    }
}

// ------------------------------------------------------------------ shapes
// Radix sequence of a DFT length: fours, then a two, then threes and fives
// (passes of radix 2 and 4 first keep the Stockham stride a power of two for
// as long as possible).  At most 16 passes.
bool factor_len(int n, FftPlan& P)
{
    P.n = n;
    P.nf = 0;
    P.code = 0;
    int m = n;
    auto push = [&](int r) {
        if (P.nf >= 16) return false;
        P.code |= (unsigned long long)r << (4 * P.nf);
        P.nf++;
        return true;
    };
    P.rg = 1;
    while (m % 4 == 0) { if (!push(4)) return false; m /= 4; }
    while (m % 2 == 0) { if (!push(2)) return false; m /= 2; }
    while (m % 3 == 0) { if (!push(3)) return false; m /= 3; }
    while (m % 5 == 0) { if (!push(5)) return false; m /= 5; }
    if (m > 1) {   // one direct-DFT pass over the remaining factor
        if (!push(0)) return false;
        P.rg = m;
    }
    return true;
}

// The nonlinear pass's plan of Np = 3 Nm: the factors 4, 2, 5 and then 3
// last, so that the last inverse pass is a radix-3 sweep whose butterflies
// hold both transforms' values at the same indices (the products are formed
// there); a factor above 5 leaves the plan generic (rg > 1), for which the
// pass keeps its plain sweeps.
bool factor_len_nl(int n, FftPlan& P)
{
    P.n = n;
    P.nf = 0;
    P.code = 0;
    P.rg = 1;
    int m = n;
    auto push = [&](int r) {
        if (P.nf >= 16) return false;
        P.code |= (unsigned long long)r << (4 * P.nf);
        P.nf++;
        return true;
    };
    while (m % 4 == 0) { if (!push(4)) return false; m /= 4; }
    while (m % 2 == 0) { if (!push(2)) return false; m /= 2; }
    while (m % 5 == 0) { if (!push(5)) return false; m /= 5; }
    while (m % 3 == 0) { if (!push(3)) return false; m /= 3; }
    if (m > 1) {
        if (!push(0)) return false;
        P.rg = m;
    }
    return true;
}

std::vector<double2> twiddles(int n)
{
    std::vector<double2> tw(n);
    for (int t = 0; t < n; ++t) {
        const long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)n;
        tw[t].x = (double)cosl(a);
        tw[t].y = -(double)sinl(a);
    }
    return tw;
}

// dcheb (chebyshev_fftw.F90:88-121): column-major N x N with row = point,
// column = polynomial; D = 2 d/dx, D2 = 4 d2/dx2 (the gap has width 1).
void host_dcheb(int N, std::vector<double>& t, std::vector<double>& D, std::vector<double>& D2, std::vector<double>& xp)
{
    const double pi = 4.0 * std::atan(1.0);
    xp.resize(N);
    t.assign((size_t)N * N, 0.0); D.assign((size_t)N * N, 0.0); D2.assign((size_t)N * N, 0.0);
    for (int i = 0; i < N; ++i) xp[i] = std::cos(((double)(N - 1 - i) * pi) / (double)(N - 1));
    std::vector<double> a(N), b(N), c(N);
    for (int i = 0; i < N; ++i) {
        a[0] = 1.0; a[1] = xp[i];
        b[0] = 0.0; b[1] = 1.0;
        c[0] = 0.0; c[1] = 0.0;
        for (int j = 2; j < N; ++j) {
            a[j] = 2.0 * xp[i] * a[j - 1] - a[j - 2];
            b[j] = 2.0 * a[j - 1] + 2.0 * xp[i] * b[j - 1] - b[j - 2];
            c[j] = 4.0 * b[j - 1] + 2.0 * xp[i] * c[j - 1] - c[j - 2];
        }
        for (int j = 0; j < N; ++j) {
            t[i + (size_t)N * j] = a[j];
            D[i + (size_t)N * j] = 2.0 * b[j];
            D2[i + (size_t)N * j] = (2.0 * 2.0) * c[j];
        }
    }
}

// The radial transforms of the Chebyshev module as dense N x N matrices
// (column-major, [row + N * col]), accumulated in long double and rounded
// once.  With n1 = N-1, p = n1 - i (the grid runs from r_min at i = 0):
//   E0 (chebinvtran, chebyshev_fftw.F90:157-194): value_i = fact/2 *
//      sum_k d_k c_k cos(pi k p / n1), d_k = 1 at k = 0, n1 and 2 otherwise;
//   Fw (chebtransform, :124-155): c_k = 2/((2N-2) fact) * sum_i d_q v_i
//      cos(pi q k / n1), q = n1 - i;
//   E1 (chebinvtranD1, :196-288): the backward recurrence beta_i = sum over
//      j > i, j - i odd, of c'_j c_j (c'_j = 4j, 2 n1 for j = n1), then
//      value_i = fact * (beta_0/2 + sum_{k=1}^{n1-1} beta_k cos(pi k p / n1));
//   E2 = the same evaluation after the recurrence twice (chebinvtranD2);
//   G = E1 Fw, d/dr of values (rhs_create.F90:79-85).
struct ChebMats {
    std::vector<double> E0, E1, E2, Fw, G;
};

const ChebMats& cheb_matrices(int N)
{
    static std::map<int, ChebMats> cache;   // guarded by g_mu (shape_get)
    auto it = cache.find(N);
    if (it != cache.end()) return it->second;
    typedef long double ld;
    const int n1 = N - 1;
    const ld pi = 3.14159265358979323846264338327950288L;
    const ld fact = std::sqrt((ld)2 / (ld)n1);
    auto cs = [&](long long a) { return std::cos(pi * (ld)(a % (2 * (long long)n1)) / (ld)n1); };
    const size_t NN = (size_t)N * N;
    std::vector<ld> e0(NN), fw(NN), wv(NN), b(NN, 0), e1(NN, 0), e2(NN, 0), g(NN, 0);
    for (int k = 0; k < N; ++k)
        for (int i = 0; i < N; ++i) {
            const ld dk = (k == 0 || k == n1) ? 1 : 2;
            e0[i + (size_t)N * k] = (fact / 2) * dk * cs((long long)k * (n1 - i));
            const int q = n1 - i;
            const ld dq = (q == 0 || q == n1) ? 1 : 2;
            fw[k + (size_t)N * i] = (ld)2 / ((ld)(2 * N - 2) * fact) * dq * cs((long long)q * k);
        }
    // evaluation of the recurrence output: W[i][k], beta_{n1} = 0 never contributes
    for (int k = 0; k < N; ++k)
        for (int i = 0; i < N; ++i) {
            const ld wk = (k == 0) ? (ld)0.5 : (k == n1 ? (ld)0.25 : (ld)1);
            wv[i + (size_t)N * k] = fact * wk * cs((long long)k * (n1 - i));
        }
    for (int j = 1; j < N; ++j) {
        const ld cj = (j == n1) ? (ld)(2 * n1) : (ld)(4 * j);
        for (int i = j - 1; i >= 0; i -= 2) b[i + (size_t)N * j] = cj;
    }
    auto mul = [&](const std::vector<ld>& A, const std::vector<ld>& B, std::vector<ld>& C) {
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                const ld bk = B[k + (size_t)N * j];
                if (bk == 0) continue;
                for (int i = 0; i < N; ++i) C[i + (size_t)N * j] += A[i + (size_t)N * k] * bk;
            }
    };
    mul(wv, b, e1);
    mul(e1, b, e2);
    mul(e1, fw, g);
    ChebMats M;
    auto cvt = [&](const std::vector<ld>& v, std::vector<double>& o) {
        o.resize(NN);
        for (size_t t = 0; t < NN; ++t) o[t] = (double)v[t];
    };
    cvt(e0, M.E0); cvt(e1, M.E1); cvt(e2, M.E2); cvt(fw, M.Fw); cvt(g, M.G);
    return cache.emplace(N, std::move(M)).first->second;
}

struct ShapeCtx {
    int N = 0, Nm = 0;
    Shape sh;
    void* mem = nullptr;   // one device allocation for all constants
    const double *t = nullptr, *D = nullptr, *D2 = nullptr;
    int refs = 0;
};

std::map<std::pair<int, int>, ShapeCtx*> g_shapes;
std::set<ShapeCtx*> g_live;   // every shape context, also those of a replaced eta

// ---- launch tuning: per (Nr_max, Nm_max) launch choices.  Every choice gives
// bitwise the same results (each output is summed in the same order), so a
// choice only changes the time.  g_tune holds choices set by fdns_tuning_set
// or fdns_ens_autotune; g_tune_epoch counts the changes, and a captured graph
// or a solve plan made before the last change is made again before its next
// use.
std::map<std::pair<int, int>, fdns_tuning> g_tune;
std::atomic<int> g_tune_epoch{0};

// Choices measured on MI355X (fdns_ens_autotune) for the shapes of the
// benchmark workloads; other shapes start from the defaults below.  97 x 256:
// the 64-run mixed ensemble; the served shapes (Nr 33..257 x Nm 64..768):
// 4-run ensembles of CNAB2, restarted SBDF3, ARS443 and BPR353 at dt = 1e-7
// (falcon-dns_amd/tune_shapes.py prints these lines).
const struct { int N, Nm; fdns_tuning t; } k_tuned[] = {
    {97, 256, {256, {8, 8, 4, 4, 4, 4}, {1, 1, 0, 0, 1, 0}, {0, 0, 0, 0, 0, 0}, 0, 0}},
    {33, 64, {256, {4, 4, 4, 8, 4, 4}, {0, 0, 0, 0, 1, 0}, {1, 1, 0, 1, 1, 1}, 8, 1}},
    {33, 128, {256, {4, 4, 4, 8, 4, 4}, {0, 0, 0, 0, 1, 0}, {1, 1, 1, 1, 1, 1}, 8, 1}},
    {33, 256, {256, {4, 4, 8, 8, 4, 4}, {0, 0, 0, 0, 1, 0}, {0, 1, 1, 1, 1, 0}, 16, 1}},
    {33, 512, {256, {4, 4, 4, 8, 8, 4}, {0, 0, 0, 0, 1, 0}, {1, 0, 0, 1, 1, 0}, 0, 1}},
    {33, 768, {256, {4, 8, 4, 8, 4, 4}, {0, 0, 0, 0, 1, 0}, {0, 1, 0, 1, 1, 0}, 0, 1}},
    {65, 64, {256, {8, 8, 4, 4, 8, 4}, {0, 1, 0, 0, 1, 0}, {0, 1, 0, 1, 1, 1}, 0, 1}},
    {65, 128, {256, {8, 8, 4, 4, 8, 4}, {0, 0, 0, 0, 1, 0}, {0, 0, 0, 1, 1, 0}, 8, 0}},
    {65, 256, {256, {8, 8, 4, 4, 8, 4}, {0, 0, 0, 0, 1, 0}, {0, 0, 0, 1, 1, 1}, 0, 1}},
    {65, 512, {256, {8, 8, 4, 4, 8, 4}, {0, 0, 0, 0, 1, 0}, {0, 0, 0, 1, 0, 0}, 0, 1}},
    {65, 768, {256, {8, 8, 4, 4, 8, 4}, {0, 1, 0, 0, 1, 0}, {0, 0, 0, 1, 0, 1}, 0, 1}},
    {129, 64, {256, {8, 8, 4, 4, 8, 4}, {0, 1, 0, 0, 1, 0}, {0, 0, 0, 1, 1, 1}, 0, 1}},
    {129, 128, {256, {8, 8, 4, 4, 8, 4}, {0, 1, 0, 0, 1, 0}, {0, 0, 0, 1, 1, 1}, 16, 0}},
    {129, 256, {256, {8, 8, 4, 4, 8, 4}, {0, 0, 0, 0, 1, 0}, {1, 1, 0, 1, 1, 0}, 16, 1}},
    {129, 512, {256, {8, 8, 4, 4, 8, 4}, {0, 0, 0, 0, 1, 0}, {0, 0, 0, 1, 0, 1}, 0, 1}},
    {129, 768, {256, {8, 8, 4, 4, 8, 4}, {0, 0, 0, 0, 1, 0}, {1, 0, 0, 1, 0, 1}, 8, 1}},
    {257, 64, {256, {8, 8, 4, 4, 8, 4}, {1, 0, 0, 0, 1, 0}, {0, 0, 1, 1, 1, 0}, 8, 0}},
    {257, 128, {256, {8, 8, 4, 4, 8, 4}, {1, 0, 0, 0, 1, 0}, {1, 0, 1, 1, 1, 1}, 8, 1}},
    {257, 256, {256, {8, 8, 4, 4, 8, 4}, {1, 0, 0, 0, 1, 0}, {1, 0, 1, 1, 1, 1}, 8, 0}},
    {257, 512, {256, {8, 8, 4, 4, 8, 4}, {1, 0, 0, 0, 1, 0}, {1, 0, 1, 1, 1, 0}, 8, 0}},
    {257, 768, {256, {8, 8, 4, 4, 8, 4}, {1, 0, 0, 0, 1, 0}, {0, 0, 1, 1, 1, 1}, 0, 1}},
};

fdns_tuning tuning_plain()
{
    fdns_tuning t;
    t.nl_tpb = 256;
    for (int c = 0; c < 6; ++c) { t.tile_waves[c] = 8; t.tile_stage[c] = 0; t.tile_pf[c] = 0; }
    t.solve_chunk = 0;
    t.lane_split = 0;
    return t;
}

fdns_tuning tuning_default(int N, int Nm)
{
    for (const auto& e : k_tuned)
        if (e.N == N && e.Nm == Nm) return e.t;
    return tuning_plain();
}

void tuning_apply(Shape& sh, const fdns_tuning& t)
{
    sh.nl_tpb = t.nl_tpb;
    for (int c = 0; c < 6; ++c) {
        sh.tile_waves[c] = t.tile_waves[c];
        sh.tile_stage[c] = t.tile_stage[c];
        sh.tile_pf[c] = t.tile_pf[c];
    }
    sh.solve_chunk = t.solve_chunk;
    sh.lane_split = t.lane_split;
}

fdns_tuning tuning_of(const Shape& sh)
{
    fdns_tuning t;
    t.nl_tpb = sh.nl_tpb;
    for (int c = 0; c < 6; ++c) {
        t.tile_waves[c] = sh.tile_waves[c];
        t.tile_stage[c] = sh.tile_stage[c];
        t.tile_pf[c] = sh.tile_pf[c];
    }
    t.solve_chunk = sh.solve_chunk;
    t.lane_split = sh.lane_split;
    return t;
}

bool tuning_valid(const fdns_tuning& t)
{
    bool ok = (t.nl_tpb == 128 || t.nl_tpb == 256) && t.solve_chunk >= 0 && t.solve_chunk <= SOLVE_RHS &&
              t.solve_chunk % 8 == 0 && (t.lane_split == 0 || t.lane_split == 1);
    for (int c = 0; c < 6; ++c)
        ok = ok && (t.tile_waves[c] == 4 || t.tile_waves[c] == 8) && (t.tile_stage[c] == 0 || t.tile_stage[c] == 1) &&
             (t.tile_pf[c] == 0 || t.tile_pf[c] == 1);
    return ok;
}

int shape_get(int N, int Nm, double eta, ShapeCtx** out)
{
    // eta enters only through r; shapes are keyed by (N, Nm, eta bits)
    std::lock_guard<std::mutex> lk(g_mu);
    (void)eta;
    auto key = std::make_pair(N, Nm);
    auto it = g_shapes.find(key);
    if (it != g_shapes.end() && it->second->sh.eta == eta) {
        it->second->refs++;
        *out = it->second;
        return 0;
    }
    if (N < 4 || N > MAX_N) return fail(FDNS_ESIZE, "Nr_max must be in [4, 513]");
    if (Nm < 1 || 3 * Nm > MAX_NP) return fail(FDNS_ESIZE, "Nm_max must be in [1, 1536]");
    auto* C = new ShapeCtx();
    C->N = N; C->Nm = Nm;
    Shape& sh = C->sh;
    sh.N = N; sh.Nm = Nm; sh.Nm1 = Nm + 1; sh.Np = 3 * Nm; sh.M = 2 * N - 2;
    sh.fact = std::sqrt(2.0 / (double)(N - 1));
    sh.eta = eta;
    {
        const auto tk = g_tune.find(key);
        tuning_apply(sh, tk != g_tune.end() ? tk->second : tuning_default(N, Nm));
    }
    if (!factor_len(sh.M, sh.pm)) { delete C; return fail(FDNS_ESIZE, "2*Nr_max-2 has too many factors"); }
    if (!factor_len(sh.Np, sh.pp) || !factor_len_nl(sh.Np, sh.pq)) {
        delete C;
        return fail(FDNS_ESIZE, "3*Nm_max has too many factors");
    }
    std::vector<double> t, D, D2, xp;
    host_dcheb(N, t, D, D2, xp);
    const double pi = 4.0 * std::atan(1.0);
    const double rmin = eta / (1.0 - eta);
    std::vector<double> rad(N), rr(N), rr2(N);
    for (int i = 0; i < N; ++i) {   // init_grid, init.F90:129-137
        rad[i] = rmin + (0.5 * (1.0 + std::cos(((double)(N - 1 - i) * pi) / (double)(N - 1))));
        rr[i] = 1.0 / rad[i];
        rr2[i] = rr[i] * rr[i];
    }
    const auto twm = twiddles(sh.M), twp = twiddles(sh.Np);
    const ChebMats& cm = cheb_matrices(N);
    const size_t nb = sizeof(double2) * (twm.size() + twp.size()) + sizeof(double) * (3 * (size_t)N + 8 * (size_t)N * N);
    if (hipMalloc(&C->mem, nb) != hipSuccess) { delete C; return fail(FDNS_ENOMEM, "hipMalloc shape constants"); }
    char* p = (char*)C->mem;
    auto put = [&](const void* src, size_t bytes) { void* d = p; hipMemcpy(d, src, bytes, hipMemcpyHostToDevice); p += bytes; return d; };
    sh.pm.tw = (const double2*)put(twm.data(), sizeof(double2) * twm.size());
    sh.pp.tw = (const double2*)put(twp.data(), sizeof(double2) * twp.size());
    sh.pq.tw = sh.pp.tw;
    sh.rad = (const double*)put(rad.data(), sizeof(double) * N);
    sh.rr = (const double*)put(rr.data(), sizeof(double) * N);
    sh.rr2 = (const double*)put(rr2.data(), sizeof(double) * N);
    C->t = (const double*)put(t.data(), sizeof(double) * t.size());
    C->D = (const double*)put(D.data(), sizeof(double) * D.size());
    C->D2 = (const double*)put(D2.data(), sizeof(double) * D2.size());
    const size_t mb = sizeof(double) * (size_t)N * N;
    sh.E0 = (const double*)put(cm.E0.data(), mb);
    sh.E1 = (const double*)put(cm.E1.data(), mb);
    sh.E2 = (const double*)put(cm.E2.data(), mb);
    sh.Fw = (const double*)put(cm.Fw.data(), mb);
    sh.G = (const double*)put(cm.G.data(), mb);
    C->refs = 1;
    if (it != g_shapes.end()) g_shapes.erase(it);   // a different eta: the old ctx lives on via its refs
    g_shapes[key] = C;
    g_live.insert(C);
    *out = C;
    return 0;
}

void shape_put(ShapeCtx* C)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (--C->refs > 0) return;
    for (auto it = g_shapes.begin(); it != g_shapes.end(); ++it)
        if (it->second == C) { g_shapes.erase(it); break; }
    g_live.erase(C);
    hipFree(C->mem);
    delete C;
}

// ------------------------------------------------------------------ operators
// The implicit operators are built once on the device (assembly, dgetrf) and
// shared by every run with the same key.  A set holds, per the run's solver,
// the packed LU factors with their row permutations (LU mode) or the inverses
// formed from them by dgetrs on unit vectors (inverse mode).  What each matrix
// depends on (mat_assembly.F90:114-150, 287-345, 393-406):
//   T sets: AT of every mode -- N, Nm, eta, wt_lhs*dt and Pr;
//   F sets: AF of modes >= 1 and A_uphi -- N, Nm, eta, wt_lhs*dt, mBC and, for
//           the LU factors only, the IMEX or IMEXRK boundary-row order (the
//           stored inverse columns are the same for both orders).
enum OpKind : int { OP_T = 0, OP_F = 1, OP_TL = 2, OP_FL = 3, OP_LAP = 4 };

struct OpKey {
    int kind, N, Nm, mBC, irk;
    uint64_t wtdt_bits, pr_bits;
    double eta;
    bool operator<(const OpKey& o) const
    {
        return std::tie(kind, N, Nm, mBC, irk, wtdt_bits, pr_bits, eta) <
               std::tie(o.kind, o.N, o.Nm, o.mBC, o.irk, o.wtdt_bits, o.pr_bits, o.eta);
    }
    bool operator==(const OpKey& o) const { return !(*this < o) && !(o < *this); }
};

struct OpSet {
    OpKey key;
    double* inv = nullptr;
    int* perm = nullptr;       // net row permutations of the LU factors held
    size_t bytes = 0;
    int refs = 0;
};

std::map<OpKey, OpSet*> g_ops;

OpKey op_key(int kind, int N, int Nm, double eta, double wtdt, double Pr, int mBC, int irk)
{
    OpKey k;
    const bool tset = (kind == OP_T || kind == OP_TL || kind == OP_LAP);   // no mBC dependence
    k.kind = kind; k.N = N; k.Nm = Nm; k.eta = eta;
    k.mBC = tset ? -1 : mBC;
    k.irk = (kind == OP_FL) ? irk : -1;
    std::memcpy(&k.wtdt_bits, &wtdt, 8);
    const double pr = tset ? Pr : 0.0;
    std::memcpy(&k.pr_bits, &pr, 8);
    return k;
}

// Doubles, then ints, of one set.
void opset_layout(int kind, int N, int Nm, size_t& nd, size_t& ni)
{
    const size_t NN = (size_t)N * N, n = (size_t)N, nm = (size_t)Nm;
    switch (kind) {
        case OP_T: nd = NN * (nm + 1); ni = n; break;                         // m = 0 keeps its LU
        case OP_F: nd = 2 * n * (n - 2) * nm + NN; ni = 0; break;
        case OP_TL: nd = NN * (nm + 1); ni = n * (nm + 1); break;
        case OP_LAP: nd = NN * (nm + 1); ni = n * (nm + 1); break;            // LAPpsi LU, every mode
        default: nd = 4 * NN * nm + NN; ni = 2 * n * nm + n; break;           // OP_FL
    }
}

size_t opset_bytes(int kind, int N, int Nm)
{
    size_t nd, ni;
    opset_layout(kind, N, Nm, nd, ni);
    return sizeof(double) * nd + sizeof(int) * ni;
}

// Device workspace for the LU factors of one chunk of modes during set-up.
constexpr size_t OP_WORKSPACE = size_t(2) << 30;

int opset_get(const ShapeCtx* C, const OpKey& k, double wtdt, double Pr, int mBC, hipStream_t st, OpSet** out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_ops.find(k);
    if (it != g_ops.end()) {
        it->second->refs++;
        *out = it->second;
        return 0;
    }
    const int N = C->N, Nm = C->Nm, Nm1 = Nm + 1;
    const size_t NN = (size_t)N * N;
    auto* S = new OpSet();
    S->key = k;
    S->bytes = opset_bytes(k.kind, N, Nm);
    if (hipMalloc((void**)&S->inv, S->bytes) != hipSuccess) { delete S; return fail(FDNS_ENOMEM, "hipMalloc operator set"); }
    size_t nd, ni;
    opset_layout(k.kind, N, Nm, nd, ni);
    S->perm = ni ? (int*)((char*)S->inv + sizeof(double) * nd) : nullptr;
    if (k.kind == OP_TL || k.kind == OP_FL || k.kind == OP_LAP) {
        // LU factors in place, every mode at once (mat_assembly.F90 + dgetrf)
        if (k.kind == OP_LAP) {
            launch_asm_LAP(S->inv, C->t, C->D, C->D2, C->sh, 0, Nm1, st);
            launch_lu(S->inv, S->perm, N, NN, Nm1, st);
        } else if (k.kind == OP_TL) {
            launch_asm_T(S->inv, C->t, C->D, C->D2, C->sh, 0, Nm1, wtdt, Pr, st);
            launch_lu(S->inv, S->perm, N, NN, Nm1, st);
        } else {
            launch_asm_F(S->inv, C->t, C->D, C->D2, C->sh, 1, Nm, wtdt, mBC, k.irk, st);
            launch_lu(S->inv, S->perm, 2 * N, 4 * NN, Nm, st);
            double* Au = S->inv + 4 * NN * (size_t)Nm;
            launch_asm_U(Au, C->t, C->D, C->D2, C->sh, wtdt, mBC, st);
            launch_lu(Au, S->perm + 2 * (size_t)N * Nm, N, NN, 1, st);
        }
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
            hipFree(S->inv);
            delete S;
            return fail(FDNS_EHIP, "operator set-up kernels failed");
        }
        S->refs = 1;
        g_ops[k] = S;
        *out = S;
        return 0;
    }
    const int n = (k.kind == OP_T) ? N : 2 * N;              // matrix order
    const size_t mat = (size_t)n * n;
    const int nmodes = (k.kind == OP_T) ? Nm1 : Nm;
    int chunk = (int)std::min<size_t>((size_t)nmodes, std::max<size_t>(1, OP_WORKSPACE / (sizeof(double) * mat + sizeof(int) * n)));
    double* work = nullptr;
    int* perm = nullptr;
    if (hipMalloc((void**)&work, sizeof(double) * mat * chunk) != hipSuccess ||
        hipMalloc((void**)&perm, sizeof(int) * (size_t)n * chunk) != hipSuccess) {
        if (work) hipFree(work);
        hipFree(S->inv);
        delete S;
        return fail(FDNS_ENOMEM, "hipMalloc operator workspace");
    }
    if (k.kind == OP_T) {
        // m = 0: the packed LU factor and its permutation, solved by dgetrs'
        // sweep (the Nusselt number is this mode's wall derivative)
        launch_asm_T(S->inv, C->t, C->D, C->D2, C->sh, 0, 1, wtdt, Pr, st);
        launch_lu(S->inv, S->perm, N, NN, 1, st);
        // m >= 1: AT^{-1}, all N columns (mat_assembly.F90:287-301)
        for (int m0 = 1; m0 < Nm1; m0 += chunk) {
            const int cm = std::min(chunk, Nm1 - m0);
            launch_asm_T(work, C->t, C->D, C->D2, C->sh, m0, cm, wtdt, Pr, st);
            launch_lu(work, perm, N, NN, cm, st);
            launch_inv(work, perm, N, NN, cm, 0, N, S->inv + (size_t)m0 * NN, NN, st);
        }
    } else {
        // columns N+1..2N-2 of AF^{-1}, modes 1..Nm (mat_assembly.F90:304-345)
        const size_t blk = 2 * (size_t)N * (size_t)(N - 2);
        for (int m0 = 1; m0 <= Nm; m0 += chunk) {
            const int cm = std::min(chunk, Nm + 1 - m0);
            launch_asm_F(work, C->t, C->D, C->D2, C->sh, m0, cm, wtdt, mBC, 0, st);
            launch_lu(work, perm, 2 * N, 4 * NN, cm, st);
            launch_inv(work, perm, 2 * N, 4 * NN, cm, N + 1, N - 2, S->inv + (size_t)(m0 - 1) * blk, blk, st);
        }
        // A_uphi^{-1} (mat_assembly.F90:393-406)
        launch_asm_U(work, C->t, C->D, C->D2, C->sh, wtdt, mBC, st);
        launch_lu(work, perm, N, NN, 1, st);
        launch_inv(work, perm, N, NN, 1, 0, N, S->inv + (size_t)Nm * blk, NN, st);
    }
    const bool ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
    hipFree(work);
    hipFree(perm);
    if (!ok) { hipFree(S->inv); delete S; return fail(FDNS_EHIP, "operator set-up kernels failed"); }
    S->refs = 1;
    g_ops[k] = S;
    *out = S;
    return 0;
}

void opset_put(OpSet* S)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (--S->refs > 0) return;
    g_ops.erase(S->key);
    hipFree(S->inv);
    delete S;
}

bool g_kernels_ready = false;

int ensure_kernels()
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_kernels_ready) return 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(FDNS_EHIP, "no HIP device");
    if (kernels_init() != 0) return fail(FDNS_EHIP, "hipFuncSetAttribute failed");
    g_kernels_ready = true;
    return 0;
}

// Device solve plan of a group of runs (or of one run stepped alone).
struct SolveSched {
    SolvePlan pl{};
    void* mem = nullptr;
    int lu = 0;
    void release()
    {
        if (mem) hipFree(mem);
        mem = nullptr;
        pl = SolvePlan{};
    }
};

}  // namespace

// ------------------------------------------------------------------ runs
// A side stream for the m = 0 temperature solve of inverse mode (k_m0, one
// workgroup per run, latency-bound): forked after the right-hand sides, it
// runs beside k_solve<0> and joins before the next kernel.
struct Side {
    hipStream_t s = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    int create()
    {
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&a, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&b, hipEventDisableTiming) != hipSuccess)
            return fail(FDNS_EHIP, "side stream");
        return 0;
    }
    void destroy()
    {
        if (s) hipStreamDestroy(s);
        if (a) hipEventDestroy(a);
        if (b) hipEventDestroy(b);
        s = nullptr;
        a = b = nullptr;
    }
};

struct fdns_run {
    fdns_params p;
    Scheme sc;
    ShapeCtx* shape = nullptr;
    OpSet* opT = nullptr;      // AT^{-1}
    OpSet* opF = nullptr;      // AF^{-1} columns, A_uphi^{-1}
    OpSet* opL = nullptr;      // LAPpsi LU factors (IMEXRK schemes with Assembly_stage)
    OpSet* stT[MAXS] = {};     // diag_diff: the operators of stage s >= 1 (own diagonal)
    OpSet* stF[MAXS] = {};
    double* lagw = nullptr;    // their Lagrange wall weights (device)
    RunDesc h;                 // host mirror (pointers and constants)
    Side side;                 // k_m0 beside the temperature solves
    RunDesc* dev = nullptr;    // home of the device descriptor
    RunDesc* own_dev = nullptr;
    void* mem = nullptr;
    size_t bytes = 0;
    hipStream_t st = nullptr;
    int kstep = 0;             // host mirror of completed steps
    bool has_state = false;
    bool in_ens = false;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int tune_epoch = 0;        // launch choices the captured graph was made with
    double* series = nullptr;
    int series_cap = 0;
    double* d_out6 = nullptr;
    SolveSched solo;           // solve plan of the run stepped alone
    double dt_new = 0.0;       // dt of the last step
    double dta[8] = {0};       // dt_array (variable dt), dta[0] the newest
    bool dta_restored = false; // dta holds a restart's checkpoint values
};

namespace {

size_t run_state_bytes(const fdns_params* p, const Scheme& sc)
{
    const size_t S = (size_t)(p->Nm + 1) * p->Nr;
    size_t nf = 11 + 4;   // fields, products, and the two solve areas (2N per mode each)
    if (sc.kind == SK_IMEX) {
        const int L = std::max(sc.n_imp, sc.n_exp);
        nf += 4 * (size_t)L;
    } else if (sc.kind == SK_RK) {
        nf += 2 * (size_t)sc.S + 2;   // FdT, FdW; upC, d1u
    } else {
        nf += 5 * (size_t)sc.S;
    }
    const int L = std::max(sc.n_imp, sc.n_exp);
    const size_t keep = p->n_KE > 0 ? S + (size_t)p->Nr : 0;   // tpass, up0prev
    return sizeof(double2) * (nf * S + (4 + 2 * (size_t)L) * p->Nr + keep) + sizeof(double) * 3 * (size_t)p->Nr +
           6 * sizeof(double);
}

int validate(const fdns_params* p, Scheme& sc)
{
    if (!p) return fail(FDNS_EINVAL, "null params");
    if (!(p->eta > 0.0 && p->eta < 1.0)) return fail(FDNS_EINVAL, "eta must be in (0,1)");
    if (!(p->dt > 0.0) || !(p->Pr > 0.0)) return fail(FDNS_EINVAL, "dt and Pr must be positive");
    if (std::strcmp(p->mBC, "NS") && std::strcmp(p->mBC, "SF")) return fail(FDNS_EINVAL, "mBC must be NS or SF");
    if (std::strcmp(p->buo_tscheme, "IMP") && std::strcmp(p->buo_tscheme, "EXP"))
        return fail(FDNS_EINVAL, "buo_tscheme must be IMP or EXP");
    int rc = scheme_lookup(p->time_scheme_type, p->time_scheme_imp, p->time_scheme_exp, sc);
    if (rc) return fail(rc, std::string("scheme not available: ") + p->time_scheme_type + "/" + p->time_scheme_imp + "/" + p->time_scheme_exp);
    if (p->Nr < 4 || p->Nr > MAX_N) return fail(FDNS_ESIZE, "Nr_max must be in [4, 513]");
    if (p->Nm < 1 || 3 * p->Nm > MAX_NP) return fail(FDNS_ESIZE, "Nm_max must be in [1, 1536]");
    FftPlan a, b;
    if (!factor_len(2 * p->Nr - 2, a)) return fail(FDNS_ESIZE, "2*Nr_max-2 has too many factors");
    if (!factor_len(3 * p->Nm, b)) return fail(FDNS_ESIZE, "3*Nm_max has too many factors");
    if (p->solver < FDNS_SOLVE_AUTO || p->solver > FDNS_SOLVE_INVERSE) return fail(FDNS_EINVAL, "unknown solver");
    if (p->l_vartimestep && !(p->CFL > 0.0 && p->dt_coef > 0.0 && p->dt_max > 0.0))
        return fail(FDNS_EINVAL, "variable dt needs CFL, dt_coef and dt_max > 0");
    if (sc.diag_diff && p->solver == FDNS_SOLVE_INVERSE)
        return fail(FDNS_EINVAL, "schemes that solve each stage with its own diagonal take LU solves only");
    if (sc.diag_diff)
        for (int k = 1; k < sc.S; ++k)
            if (sc.aD[k * MAXS + k] == 0.0)
                // mat_build_IRK with wt = 0: omega's interior rows become
                // omega(r_i) = rhs and nothing else involves omega's wall
                // values, so the vorticity-streamfunction matrix is singular
                return fail(FDNS_ESCHEME, "a stage with a zero diagonal (aD(k,k) = 0, k > 1) makes the "
                                          "vorticity-streamfunction operator singular");
    if (((sc.kind == SK_IMEXRK && !sc.sa) || sc.kind == SK_RK) &&
        (p->lagpts < 2 || p->lagpts > LAGMAX || p->lagpts > p->Nr - 2))
        return fail(FDNS_EINVAL, "lagpts must be in [2, min(64, Nr_max-2)] for this scheme");
    return 0;
}

// LU mode (dgetrs order) or inverse mode for these parameters.
bool solver_lu(const fdns_params* p)
{
    if (p->solver == FDNS_SOLVE_LU) return true;
    if (p->solver == FDNS_SOLVE_INVERSE) return false;
    if (p->l_vartimestep) return true;   // re-factored whenever dt changes: no inverses
    Scheme sc;
    if (scheme_lookup(p->time_scheme_type, p->time_scheme_imp, p->time_scheme_exp, sc) == 0 && sc.diag_diff)
        return true;                     // a factor set per stage: solved by LU
    return p->Nr > 129;
}

// wt_lhs*dt of mat_build (IMEX) or butcher_aD(diag,diag)*dt of mat_build_IRK
// (fourierloop_imexrk.F90:109-117), and the keys of the run's two operator sets.
void keys_for(const fdns_params* p, const Scheme& sc, double wtdt, OpKey& kT, OpKey& kF)
{
    const int mBC = std::strcmp(p->mBC, "NS") ? 1 : 0;
    const int irk = sc.kind == SK_IMEXRK ? 1 : 0;
    const bool lu = solver_lu(p);
    kT = op_key(lu ? OP_TL : OP_T, p->Nr, p->Nm, p->eta, wtdt, p->Pr, mBC, irk);
    kF = op_key(lu ? OP_FL : OP_F, p->Nr, p->Nm, p->eta, wtdt, p->Pr, mBC, irk);
}

void run_keys(const fdns_params* p, const Scheme& sc, OpKey& kT, OpKey& kF, double* wtdt_out = nullptr)
{
    double wl = 0.0, wi[MAXH], we[MAXH];
    const double dta[4] = {p->dt, p->dt, p->dt, p->dt};
    if (sc.kind == SK_IMEX) imex_weights(sc, dta, wl, wi, we);
    else wl = sc.aD[sc.diag * MAXS + sc.diag];
    const double wtdt = wl * p->dt;
    keys_for(p, sc, wtdt, kT, kF);
    if (wtdt_out) *wtdt_out = wtdt;
}

// dlagrange / d2lagrange (chebyshev_fftw.F90:449-517): first and second
// derivative weights at x of the Lagrange interpolant through xp[0..n).
void lagrange_weights(int n, const double* xp, double x, double* dw, double* d2w)
{
    for (int j = 0; j < n; ++j) {
        dw[j] = 0.0;
        d2w[j] = 0.0;
    }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            if (i == j) continue;
            double w = 1.0;
            for (int m = 0; m < n; ++m)
                if (m != i && m != j) w = w * (x - xp[m]) / (xp[j] - xp[m]);
            dw[j] = dw[j] + w / (xp[j] - xp[i]);
        }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            if (i == j) continue;
            double mw = 0.0;
            for (int m = 0; m < n; ++m) {
                if (m == i || m == j) continue;
                double w = 1.0;
                for (int l = 0; l < n; ++l)
                    if (l != j && l != i && l != m) w = w * (x - xp[l]) / (xp[j] - xp[l]);
                mw = mw + w / (xp[j] - xp[m]);
            }
            d2w[j] = d2w[j] + mw / (xp[j] - xp[i]);
        }
}

// The device pointers of the run's two operator sets.
void bind_ops(fdns_run* R)
{
    RunDesc& h = R->h;
    const size_t N = (size_t)R->p.Nr, Nm = (size_t)R->p.Nm;
    if (!R->opT || !R->opF) {   // the explicit RK path: LAPpsi only
        if (R->opL) {
            h.Li = R->opL->inv;
            h.pL = R->opL->perm;
        }
        return;
    }
    h.Ti = R->opT->inv;
    h.pT = R->opT->perm;
    h.Fi = R->opF->inv;
    if (h.solve_lu) {
        h.pF = R->opF->perm;
        h.Ui = R->opF->inv + 4 * N * N * Nm;
        h.pU = R->opF->perm + 2 * N * Nm;
    } else {
        h.Ui = R->opF->inv + 2 * N * (N - 2) * Nm;
    }
    if (R->opL) {
        h.Li = R->opL->inv;
        h.pL = R->opL->perm;
    }
    for (int k = 0; k < MAXS; ++k) {
        const OpSet* T = (R->sc.diag_diff && R->stT[k]) ? R->stT[k] : R->opT;
        const OpSet* F = (R->sc.diag_diff && R->stF[k]) ? R->stF[k] : R->opF;
        h.TiS[k] = T->inv;
        h.pTS[k] = T->perm;
        h.FiS[k] = F->inv;
        if (h.solve_lu) {
            h.pFS[k] = F->perm;
            h.UiS[k] = F->inv + 4 * N * N * Nm;
            h.pUS[k] = F->perm + 2 * N * Nm;
        } else {
            h.UiS[k] = F->inv + 2 * N * (N - 2) * Nm;
        }
    }
}

// diag_diff: the operator sets of stages 1..S-1, mat_build_IRK with the
// stage's own diagonal (fourierloop_imexrk.F90:134-158); stages with equal
// diagonals share a set through the cache.
int stage_ops(fdns_run* R, double dt)
{
    if (!R->sc.diag_diff) return 0;
    const int mBC = std::strcmp(R->p.mBC, "NS") ? 1 : 0;
    for (int k = 1; k < R->sc.S; ++k) {
        const double wtdt = R->sc.aD[k * MAXS + k] * dt;
        OpKey kT, kF;
        keys_for(&R->p, R->sc, wtdt, kT, kF);
        OpSet *nT = nullptr, *nF = nullptr;
        int rc = opset_get(R->shape, kT, wtdt, R->p.Pr, mBC, R->st, &nT);
        if (!rc) rc = opset_get(R->shape, kF, wtdt, R->p.Pr, mBC, R->st, &nF);
        if (rc) {
            if (nT) opset_put(nT);
            return rc;
        }
        if (R->stT[k]) opset_put(R->stT[k]);
        if (R->stF[k]) opset_put(R->stF[k]);
        R->stT[k] = nT;
        R->stF[k] = nF;
    }
    return 0;
}

int run_zero_dynamic(fdns_run* R)
{
    // everything except the factors and constants starts at zero: fields,
    // products, histories, stage terms (init.F90:164-275, timeschemes.F90:51-64)
    HIPCHK(hipMemsetAsync(R->mem, 0, R->bytes, R->st));
    return 0;
}

int upload_desc(fdns_run* R)
{
    HIPCHK(hipMemcpyAsync(R->dev, &R->h, sizeof(RunDesc), hipMemcpyHostToDevice, R->st));
    return 0;
}

int download_desc(fdns_run* R)
{
    RunDesc tmp;
    HIPCHK(hipMemcpyAsync(&tmp, R->dev, sizeof(RunDesc), hipMemcpyDeviceToHost, R->st));
    HIPCHK(hipStreamSynchronize(R->st));
    R->h.head = tmp.head; R->h.kstep = tmp.kstep; R->h.time = tmp.time; R->h.nrec = tmp.nrec;
    return 0;
}

// Runs per batched solve chunk: the staged right-hand sides of a chunk
// (K4 x 16 ceil(2 n / 16) doubles) stay within 80 KiB of LDS, so that two
// solve workgroups share a CU at every size.
int solve_chunk_cap(int N)
{
    const size_t k4 = 4 * (((size_t)N + 3) / 4);
    int cap = SOLVE_RHS;
    while (cap > 8 && sizeof(double) * k4 * (size_t)(16 * ((2 * cap + 15) / 16)) > 80 * 1024) cap -= 8;
    return cap;
}

// Solve plan of a group: the runs that share an operator set (the same OpSet)
// in chunks of at most solve_chunk_cap right-hand sides, separately for the
// temperature and the vorticity / mean-flow sets.
int plan_build(const std::vector<fdns_run*>& runs, SolveSched& S)
{
    S.release();
    S.lu = runs[0]->h.solve_lu;
    const int want = runs[0]->shape->sh.solve_chunk;
    const int cap = want > 0 ? std::min(want, solve_chunk_cap(runs[0]->p.Nr)) : solve_chunk_cap(runs[0]->p.Nr);
    std::vector<int> idx[2];
    std::vector<int2> ch[2];
    int wmax[2] = {0, 0};
    for (int t = 0; t < 2; ++t) {
        std::vector<const OpSet*> sets;
        std::vector<std::vector<int>> members;
        for (size_t i = 0; i < runs.size(); ++i) {
            const OpSet* o = t == 0 ? runs[i]->opT : runs[i]->opF;
            size_t k = 0;
            while (k < sets.size() && sets[k] != o) ++k;
            if (k == sets.size()) {
                sets.push_back(o);
                members.emplace_back();
            }
            members[k].push_back((int)i);
        }
        for (const auto& mem : members)
            for (size_t a = 0; a < mem.size(); a += (size_t)cap) {
                const int n = (int)std::min<size_t>((size_t)cap, mem.size() - a);
                ch[t].push_back(make_int2((int)idx[t].size(), n));
                idx[t].insert(idx[t].end(), mem.begin() + a, mem.begin() + a + n);
                wmax[t] = std::max(wmax[t], 16 * ((2 * n + 15) / 16));
            }
    }
    const size_t nc = ch[0].size() + ch[1].size(), ni = idx[0].size() + idx[1].size();
    std::vector<char> host(sizeof(int2) * nc + sizeof(int) * ni);
    std::memcpy(host.data(), ch[0].data(), sizeof(int2) * ch[0].size());
    std::memcpy(host.data() + sizeof(int2) * ch[0].size(), ch[1].data(), sizeof(int2) * ch[1].size());
    std::memcpy(host.data() + sizeof(int2) * nc, idx[0].data(), sizeof(int) * idx[0].size());
    std::memcpy(host.data() + sizeof(int2) * nc + sizeof(int) * idx[0].size(), idx[1].data(), sizeof(int) * idx[1].size());
    if (hipMalloc(&S.mem, host.size()) != hipSuccess) return fail(FDNS_ENOMEM, "solve plan");
    HIPCHK(hipMemcpy(S.mem, host.data(), host.size(), hipMemcpyHostToDevice));
    int2* c0 = (int2*)S.mem;
    int* i0 = (int*)(c0 + nc);
    S.pl.chT = c0;
    S.pl.chF = c0 + ch[0].size();
    S.pl.idxT = i0;
    S.pl.idxF = i0 + idx[0].size();
    S.pl.nchT = (int)ch[0].size();
    S.pl.nchF = (int)ch[1].size();
    S.pl.wT = wmax[0];
    S.pl.wF = wmax[1];
    return 0;
}

// Optional per-launch timing with HIP events recorded on the launch stream,
// summed per kernel class (FDNS_KC_* of include/fdns.h).
struct KTimer {
    static constexpr int NCLS = FDNS_NKCLASS;
    std::vector<hipEvent_t> ev;
    std::vector<int> cls;
    size_t used = 0;
    void begin(hipStream_t st)
    {
        if (used + 2 > ev.size()) {
            for (int k = 0; k < 64; ++k) {
                hipEvent_t e;
                if (hipEventCreate(&e) == hipSuccess) ev.push_back(e);
            }
            cls.resize(ev.size() / 2 + 1);
        }
        hipEventRecord(ev[used], st);
    }
    void end(int c, hipStream_t st)
    {
        hipEventRecord(ev[used + 1], st);
        cls[used / 2] = c;
        used += 2;
    }
    // sums per class after the stream has drained
    void collect(double* ms, int* count)
    {
        for (int c = 0; c < NCLS; ++c) { ms[c] = 0.0; count[c] = 0; }
        for (size_t k = 0; k + 1 < used; k += 2) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, ev[k], ev[k + 1]) == hipSuccess) {
                ms[cls[k / 2]] += t;
                count[cls[k / 2]] += 1;
            }
        }
        used = 0;
    }
    ~KTimer()
    {
        for (auto e : ev) hipEventDestroy(e);
    }
};

#define TIMED(kt, c, st, call)          \
    do {                                \
        if (kt) (kt)->begin(st);        \
        call;                           \
        if (kt) (kt)->end((c), (st));   \
    } while (0)

// The m = 0 temperature solves (inverse mode) and the temperature solves: for
// Nr_max <= 100 one launch whose first blocks do the m = 0 systems (a
// replayed graph's fork and join cost ~20 us per solve stage); larger, k_m0
// on the side stream beside k_solve when one is given.  Under the kernel
// timer, which times each kernel alone: k_m0, then k_solve.
void m0_and_solve_T(const RunDesc* dev, int n, const Shape& sh, const SolveSched& ss, int stage, hipStream_t st,
                    KTimer* kt, const Side* sd)
{
    if (!kt && !ss.lu && launch_solve_T_m0(dev, n, ss.pl, sh, stage, st)) return;
    const bool fork = sd && !kt && !ss.lu;
    if (fork) {
        // k_solve, the longer of the two, launched first on the owner's stream
        // (a replayed graph keeps it on the chain's queue), k_m0 forked
        hipEventRecord(sd->a, st);
        launch_solve(dev, n, ss.pl, sh, 0, ss.lu, stage, st);
        hipStreamWaitEvent(sd->s, sd->a, 0);
        launch_m0(dev, n, sh, sd->s);
        hipEventRecord(sd->b, sd->s);
        hipStreamWaitEvent(st, sd->b, 0);
        return;
    }
    if (!ss.lu) TIMED(kt, FDNS_KC_OTHER, st, launch_m0(dev, n, sh, st));
    TIMED(kt, FDNS_KC_SOLVE_T, st, launch_solve(dev, n, ss.pl, sh, 0, ss.lu, stage, st));
}

// The mode update of one IMEX step: right-hand sides, temperature solves,
// new temperature and vorticity right-hand sides, vorticity solves, new fields.
void imex_launches(const RunDesc* dev, int n, const Shape& sh, const SolveSched& ss, int first, hipStream_t st,
                   KTimer* kt, const Side* sd = nullptr)
{
    if (first) TIMED(kt, FDNS_KC_OTHER, st, launch_imex_prep(dev, n, sh, st));
    TIMED(kt, FDNS_KC_IMEX_PRE, st, launch_imex_pre(dev, n, sh, st));
    m0_and_solve_T(dev, n, sh, ss, 0, st, kt, sd);
    TIMED(kt, FDNS_KC_IMEX_MID, st, launch_imex_mid(dev, n, sh, st));
    TIMED(kt, FDNS_KC_SOLVE_F, st, launch_solve(dev, n, ss.pl, sh, 1, ss.lu, 0, st));
    TIMED(kt, FDNS_KC_IMEX_POST, st, launch_imex_post(dev, n, sh, st));
}

// One IMEXRK stage, in the same five kernels.
void stage_launches(const RunDesc* dev, int n, const Shape& sh, const SolveSched& ss, int s, hipStream_t st,
                    KTimer* kt, const Side* sd = nullptr)
{
    TIMED(kt, FDNS_KC_RK_PRE, st, launch_rk_pre(dev, n, sh, s, st));
    m0_and_solve_T(dev, n, sh, ss, s, st, kt, sd);
    TIMED(kt, FDNS_KC_RK_MID, st, launch_rk_mid(dev, n, sh, s, st));
    TIMED(kt, FDNS_KC_SOLVE_F, st, launch_solve(dev, n, ss.pl, sh, 1, ss.lu, s, st));
    TIMED(kt, FDNS_KC_RK_POST, st, launch_rk_post(dev, n, sh, s, st));
}

// The launch sequence of one time step of one group of runs that share a
// shape, a scheme structure and a solver mode.  first: this is the group's
// first step.
void step_launches(const RunDesc* dev, int n, const Shape& sh, const Scheme& sc, const SolveSched& ss, bool first,
                   bool restart, bool record, hipStream_t st, KTimer* kt = nullptr, bool keep = false,
                   const Side* sd = nullptr)
{
    if (sc.kind == SK_IMEX) {
        if (keep) TIMED(kt, FDNS_KC_OTHER, st, launch_keep(dev, n, sh, 1, st));   // tt, upFR_prev of the step's start
        if (!first || restart) TIMED(kt, FDNS_KC_NONLIN, st, launch_nonlin(dev, n, sh, st));   // steptime.F90:202
        imex_launches(dev, n, sh, ss, first ? 1 : 0, st, kt, sd);
    } else if (sc.kind == SK_RK) {
        // timeloop_rk (steptime_rk.F90:108-133): stage values from stage 2,
        // the nonlinear pass (not at stage 1 of a fresh start's first step),
        // the stage's right-hand sides; Assembly_stage ends the step
        for (int s = 0; s < sc.S; ++s) {
            if (s > 0) TIMED(kt, FDNS_KC_OTHER, st, launch_rkx_update(dev, n, sh, s, st));
            if (keep && s == sc.S - 1) TIMED(kt, FDNS_KC_OTHER, st, launch_keep(dev, n, sh, 0, st));
            if (!first || s > 0 || restart) TIMED(kt, FDNS_KC_NONLIN, st, launch_nonlin(dev, n, sh, st));
            TIMED(kt, FDNS_KC_OTHER, st, launch_rkx_rhs(dev, n, sh, s, st));
        }
        TIMED(kt, FDNS_KC_OTHER, st, launch_rkx_update(dev, n, sh, sc.S, st));
    } else {
        // steptime_imexrk.F90:143-188.  Stiffly accurate: the last stage is
        // the new state and its terms are the next step's first-stage terms.
        // Otherwise the first-stage terms come from the state (k_rk_init, in
        // the prologue on the first step), a nonlinear pass follows the last
        // stage when its column is used, and Assembly_stage ends the step.
        if (!sc.sa && !first) TIMED(kt, FDNS_KC_OTHER, st, launch_rk_init(dev, n, sh, st));
        // the step's passes: stage 0's, then after stage s; the temperature of
        // the last one is kept for the power budget
        auto pass = [&](int s) {
            return s == 0 ? ((!first || restart) && sc.nl_col[0]) : ((s < sc.S - 1 || !sc.sa) && sc.nl_col[s]);
        };
        int last = -1;
        for (int s = 0; s < sc.S; ++s)
            if (pass(s)) last = s;
        if (keep && last == 0) TIMED(kt, FDNS_KC_OTHER, st, launch_keep(dev, n, sh, 0, st));
        if (pass(0)) TIMED(kt, FDNS_KC_NONLIN, st, launch_nonlin(dev, n, sh, st));
        for (int s = 1; s < sc.S; ++s) {
            stage_launches(dev, n, sh, ss, s, st, kt, sd);
            if (keep && last == s) TIMED(kt, FDNS_KC_OTHER, st, launch_keep(dev, n, sh, 0, st));
            if (pass(s)) TIMED(kt, FDNS_KC_NONLIN, st, launch_nonlin(dev, n, sh, st));
        }
        if (!sc.sa) TIMED(kt, FDNS_KC_OTHER, st, launch_rk_asm(dev, n, sh, st));
    }
    TIMED(kt, FDNS_KC_OTHER, st, launch_advance((RunDesc*)dev, n, st));
    if (record) TIMED(kt, FDNS_KC_OTHER, st, launch_diag(dev, n, sh, 1, nullptr, st));
}

// Work that precedes the first step: the restart history start-up (IMEX) or
// the stage-0 terms of the initial state (IMEXRK).
void prologue_launches(const RunDesc* dev, int n, const Shape& sh, const Scheme& sc, bool restart, hipStream_t st)
{
    if (sc.kind == SK_IMEX) {
        if (restart) {
            launch_nonlin(dev, n, sh, st);
            launch_imex_restart(dev, n, sh, st);
        }
    } else if (sc.kind == SK_RK) {
        launch_rkx_init(dev, n, sh, st);
    } else {
        launch_rk_init(dev, n, sh, st);
    }
}

// ---- variable dt
// compute_new_dt (steptime.F90:313-370 for IMEX, steptime_imexrk.F90:248-308
// for IMEXRK) from the per-level velocity maxima of the last nonlinear pass:
// get_dtvals (nonlin.F90:93-106) gives dr/max|u_r| and (2/3) pi r /
// (Nm max|u_phi|) per level; their minimum, scaled by CFL (and 1/Pr for
// Pr > 1), limits dt, capped by dt_max; l_optimizedt keeps dt unless the
// limit moves past it.  ks = n_step - n_restart.  dt_array: IMEX shifts every
// slot, IMEXRK only slot 2.
void new_dt(fdns_run* R, const double* cfl, int ks, bool imexrk)
{
    const fdns_params& p = R->p;
    const int N = p.Nr;
    const double pi = 4.0 * std::atan(1.0);
    const double rmin = p.eta / (1.0 - p.eta);
    auto rad = [&](int i) { return rmin + (0.5 * (1.0 + std::cos(((double)(N - 1 - i) * pi) / (double)(N - 1)))); };
    double dt_cal = INFINITY;
    for (int i = 0; i < N; ++i) {
        const double dr = i > 0 ? rad(i) - rad(i - 1) : rad(1) - rad(0);   // dr(1) = dr(2), init.F90:133-137
        const double vr = dr / cfl[2 * i];
        const double vp = (2.0 / 3.0) * (rad(i) * pi) / ((double)p.Nm * cfl[2 * i + 1]);
        dt_cal = std::fmin(dt_cal, std::fmin(vr, vp));
    }
    const double dt_n = (p.Pr == 1.0 || p.Pr > 1.0) ? std::fmax(1.0 / p.Pr, 1.0) * p.CFL * dt_cal
                                                    : std::fmin(1.0 / p.Pr, 1.0) * p.CFL * dt_cal;
    const double dt2 = std::fmin(0.5 * (1.0 / p.dt_coef + 1.0) * dt_n, p.dt_max);
    double* a = R->dta;
    if (p.l_optimizedt) {
        if (a[1] > p.dt_max && ks > 1) R->dt_new = p.dt_max;
        else if (a[1] >= dt_cal && ks > 1) R->dt_new = dt2;
        else if (p.dt_coef * a[1] < dt_cal && a[1] < p.dt_max && ks > 1) R->dt_new = (a[0] < dt2) ? dt2 : a[0];
        // otherwise dt keeps the previous step's value
    } else {
        R->dt_new = dt2;
    }
    const int nmax = std::max(R->sc.n_imp, R->sc.n_exp);
    const bool restart = p.l_restart != 0;
    if (!restart && ks == 1) std::fill(a, a + nmax, R->dt_new);
    else if (restart && ks == 1) a[0] = R->dt_new;
    if (ks > 1) {
        if (nmax > 1) {
            if (imexrk) a[1] = a[0];
            else for (int i = nmax - 1; i >= 1; --i) a[i] = a[i - 1];
        }
        a[0] = R->dt_new;
    }
}

// compute_new_dt of the RK path (steptime_rk.F90:181-236): the CFL limits of
// the last pass and the diffusive ones dr^2 and ((2/3) pi r / Nm)^2
// (nonlin.F90:102-104); from the second step the optimize branches, with
// dt_array(1) = dt_array(2) = dt (reset at every step's start, :96-99); no
// branch taken keeps dt.
void new_dt_rk(fdns_run* R, const double* cfl, int ks)
{
    const fdns_params& p = R->p;
    const int N = p.Nr;
    const double pi = 4.0 * std::atan(1.0);
    const double rmin = p.eta / (1.0 - p.eta);
    auto rad = [&](int i) { return rmin + (0.5 * (1.0 + std::cos(((double)(N - 1 - i) * pi) / (double)(N - 1)))); };
    R->dt_new = p.dt;
    if (ks > 1) {
        double dt_cal = INFINITY;
        for (int i = 0; i < N; ++i) {
            const double dr = i > 0 ? rad(i) - rad(i - 1) : rad(1) - rad(0);
            const double vr = dr / cfl[2 * i];
            const double vp = (2.0 / 3.0) * (rad(i) * pi) / ((double)p.Nm * cfl[2 * i + 1]);
            const double rkr = dr * dr;
            const double rkp = (2.0 / 3.0) * (rad(i) * pi) / (double)p.Nm * (2.0 / 3.0) * (rad(i) * pi) / (double)p.Nm;
            dt_cal = std::fmin(dt_cal, std::fmin(std::fmin(vr, vp), std::fmin(rkr, rkp)));
        }
        const double dt_n = (p.Pr == 1.0 || p.Pr > 1.0) ? std::fmax(1.0 / p.Pr, 1.0) * p.CFL * dt_cal
                                                        : std::fmin(1.0 / p.Pr, 1.0) * p.CFL * dt_cal;
        const double dt2 = std::fmin(0.5 * (1.0 / p.dt_coef + 1.0) * dt_n, p.dt_max);
        const double d = p.dt;
        if (d > p.dt_max) R->dt_new = p.dt_max;
        else if (d >= dt_cal) R->dt_new = dt2;
        else if (p.dt_coef * d < dt_cal && d < p.dt_max) R->dt_new = (d < dt2) ? dt2 : d;
    }
    std::fill(R->dta, R->dta + 8, p.dt);
    R->dta[0] = R->dt_new;
}

// mat_build with a new wt_lhs * dt: the operator sets of that key (built, or
// shared with runs that have it), replacing the run's current ones.
int rebind_ops(fdns_run* R, double wl, double dt)
{
    const double wtdt = wl * dt;
    OpKey kT, kF;
    keys_for(&R->p, R->sc, wtdt, kT, kF);
    if (kT == R->opT->key && kF == R->opF->key) return 0;
    int rs = stage_ops(R, dt);
    if (rs) return rs;
    const int mBC = std::strcmp(R->p.mBC, "NS") ? 1 : 0;
    OpSet *nT = nullptr, *nF = nullptr;
    int rc = opset_get(R->shape, kT, wtdt, R->p.Pr, mBC, R->st, &nT);
    if (!rc) rc = opset_get(R->shape, kF, wtdt, R->p.Pr, mBC, R->st, &nF);
    if (rc) {
        if (nT) opset_put(nT);
        return rc;
    }
    opset_put(R->opT);
    opset_put(R->opF);
    R->opT = nT;
    R->opF = nF;
    bind_ops(R);
    return 0;
}

// Gaspari & Cohn's compact Gaussian-like pieces as init.F90:937-985 writes
// them: the middle part (distance r < c), the edge part (c <= r < 2c) and
// the centre value that normalises both.
double gasp_middle(double r, double c, double L)
{
    const double pi = 4.0 * std::atan(1.0);
    return pi / 3.0 * L * r * (r + 3.0 * L) * std::exp(-r / L) +
           (2.0 * pi * L * L * ((c + L) * (c + L))) / r * std::exp(-(2.0 * c) / L) *
               (1.0 - (1.0 - r / (c + L)) * std::exp(r / L)) +
           pi * L * L * L * (std::exp(-r / L) - std::exp((r - 2.0 * c) / L));
}

double gasp_edge(double r, double c, double L)
{
    const double pi = 4.0 * std::atan(1.0);
    return 2.0 * pi * L / r * std::exp(-r / L) *
           (0.5 * (r * (r + L) * (2.0 * c - r)) + 1.0 / 3.0 * ((r - c) * (r - c) * (r - c) - c * c * c) -
            L * (c + L) * (r - c + L) + L * ((c + L) * (c + L)) * std::exp((r - 2.0 * c) / L));
}

double gasp_center(double c, double L)
{
    const double pi = 4.0 * std::atan(1.0);
    return pi * L * L * L * (1.0 - std::exp(-(2.0 * c) / L)) - 2.0 * pi * c * L * (c + L) * std::exp(-(2.0 * c) / L);
}

// The physical temperature tt(Np, Nr) of init_perturbation (init.F90:277-399)
// on top of the conductive profile tcon (init.F90:207-211):
//   n_init = 0  a compact bubble at (r_c, pi): Gaspari pieces in r (sigma_r =
//               0.1/sqrt 2, L = sigma_r/2, r_c the grid point nearest the
//               mid radius) times exp(-(phi - pi)^2 / sigma_p^2), sigma_p the
//               default-real 0.2; the inner branch computed with c2 is
//               normalised by the centre value of c1 and the outer branch
//               the other way round, as the reference does;
//   n_init < 0  a Gaussian at (mid radius, phi = 2) for 1 < phi < 3, times
//               -(rmax - r)(rmin - r) ampT / (2 pi sigma_p sigma_r), sigma = 0.25;
//   n_init > 0  ampT (rmax - r)(rmin - r) cos(n_init phi).
// phi is init_grid's accumulated grid phi(i) = phi(i-1) + 2 pi/Np
// (init.F90:141-147); the bubble uses (i-1) 2 pi/Np directly.
void perturbation_field(int N, int Nm, double eta, int n_init, double ampT, std::vector<double>& tt)
{
    const int Np = 3 * Nm;
    const double pi = 4.0 * std::atan(1.0);
    const double rmin = eta / (1.0 - eta), rmax = 1.0 / (1.0 - eta);
    std::vector<double> rad(N), phi(Np), tcon(N);
    for (int i = 0; i < N; ++i) rad[i] = rmin + (0.5 * (1.0 + std::cos(((double)(N - 1 - i) * pi) / (double)(N - 1))));
    for (int j = 0; j < N; ++j) tcon[j] = std::log(rad[j]) / std::log(rmin / rmax) - std::log(rmax) / std::log(rmin / rmax);
    phi[0] = 0.0;
    for (int i = 1; i < Np; ++i) phi[i] = phi[i - 1] + 2.0 * pi / (double)Np;
    tt.assign((size_t)Np * N, 0.0);
    auto at = [&](int i, int j) -> double& { return tt[(size_t)i + (size_t)Np * j]; };
    if (n_init == 0) {
        const double sigma_r = 0.1 / std::sqrt(2.0);
        const double sigma_p = (double)0.2f;          // default-real literal
        double rc = 0.5 * (rmax + rmin);
        int ir = 0;                                   // minloc: the first nearest point
        for (int j = 1; j < N; ++j)
            if (std::fabs(rad[j] - rc) < std::fabs(rad[ir] - rc)) ir = j;
        rc = rad[ir];
        const double c1 = 0.5 * (rad[N - 1] - rc), c2 = 0.5 * (rc - rad[0]), L = 0.5 * sigma_r;
        std::vector<double> g(N, 0.0);
        for (int j = 0; j < N; ++j) {
            if (j == ir) g[j] = 1.0;
            else if (rad[j] < rc - c2) g[j] = gasp_edge(rc - rad[j], c2, L);
            else if (rad[j] >= rc - c2 && rad[j] < rc) g[j] = gasp_middle(rc - rad[j], c2, L);
            else if (rad[j] > rc && rad[j] <= rc + c1) g[j] = gasp_middle(rad[j] - rc, c1, L);
            else if (rad[j] > rc + c1) g[j] = gasp_edge(rad[j] - rc, c1, L);
        }
        for (int j = 0; j < N; ++j) {
            if (j < ir) g[j] = g[j] / gasp_center(c1, L);
            else if (j > ir) g[j] = g[j] / gasp_center(c2, L);
        }
        const double phi0 = pi;
        for (int j = 0; j < N; ++j) {
            const double c_r = ampT * g[j];
            for (int i = 0; i < Np; ++i) {
                const double ph = (double)i * 2.0 * pi / (double)Np;
                at(i, j) = c_r * std::exp(-((ph - phi0) * (ph - phi0)) / (sigma_p * sigma_p)) + tcon[j];
            }
        }
    } else if (n_init < 0) {
        const double sigma_r = 0.25, sigma_p = 0.25;
        const double rm = rmax - (rmax - rmin) / 2.0;
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < Np; ++i) {
                if (phi[i] > 1.0 && phi[i] < 3.0)
                    at(i, j) = -1.0 * (rmax - rad[j]) * (rmin - rad[j]) * ampT * (1.0 / (2.0 * pi * sigma_p * sigma_r)) *
                                   std::exp(-1.0 * ((rad[j] - rm) * (rad[j] - rm) / (sigma_r * sigma_r) +
                                                    (phi[i] - 2.0) * (phi[i] - 2.0) / (sigma_p * sigma_p))) +
                               tcon[j];
                else
                    at(i, j) = tcon[j];
            }
    } else {
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < Np; ++i)
                at(i, j) = ampT * (rmax - rad[j]) * (rmin - rad[j]) * std::cos((double)n_init * phi[i]) + tcon[j];
    }
}

// The steps of a variable-dt run, one at a time: the step's nonlinear pass
// records the velocity maxima, the host computes the new dt from them, and
// the rest of the step runs with the new weights and, when dt changed, newly
// factored operators (fourierloop.F90:86-94, fourierloop_imexrk.F90:108-118).
// steptime.F90:198-231 (IMEX) and steptime_imexrk.F90:102-138 (IMEXRK).
int vardt_steps(fdns_run* R, int nsteps)
{
    const Shape& sh = R->shape->sh;
    const bool restart = R->p.l_restart != 0, imex = R->sc.kind == SK_IMEX, rk = R->sc.kind == SK_RK;
    const int N = R->p.Nr, nKE = R->p.n_KE;
    std::vector<double> cfl(2 * (size_t)N);
    int rc;
    for (int k = 0; k < nsteps; ++k) {
        const bool first = R->kstep == 0;
        const int ks = R->kstep + 1;
        const bool rec = nKE > 0 && (ks % nKE) == 0;
        if (first) {
            prologue_launches(R->dev, 1, sh, R->sc, restart, R->st);
            R->dt_new = R->p.dt;
            // the IMEX loop keeps a restart's dt_array (steptime.F90:198-231);
            // the IMEXRK and RK loops reset it (steptime_imexrk.F90:102-106,
            // steptime_rk.F90:96-99)
            if (!(imex && R->dta_restored)) std::fill(R->dta, R->dta + 8, R->p.dt);
        }
        // IMEX: the step's own pass (none on a fresh first step); IMEXRK: an
        // extra pass on a restart's first step, else the previous step's last
        // (not stiffly accurate: a pass over the state, steptime_imexrk.F90:118-125)
        if (R->h.keep) launch_keep(R->dev, 1, sh, imex ? 1 : 0, R->st);
        // the RK path takes the previous step's last pass, with no extra one
        if (!rk && ((imex && (!first || restart)) || (!imex && first && restart) || (!imex && !R->sc.sa && !first)))
            launch_nonlin(R->dev, 1, sh, R->st);
        HIPCHK(hipGetLastError());
        if (rk) {
            if (!first) {
                HIPCHK(hipMemcpyAsync(cfl.data(), R->h.cfl.p, sizeof(double) * 2 * N, hipMemcpyDeviceToHost, R->st));
                HIPCHK(hipStreamSynchronize(R->st));
            }
            new_dt_rk(R, cfl.data(), ks);
        } else if (!first || restart) {
            HIPCHK(hipMemcpyAsync(cfl.data(), R->h.cfl.p, sizeof(double) * 2 * N, hipMemcpyDeviceToHost, R->st));
            HIPCHK(hipStreamSynchronize(R->st));
            new_dt(R, cfl.data(), ks, !imex);
        }
        if ((rc = download_desc(R))) return rc;     // head, kstep, time, records
        RunDesc& h = R->h;
        h.dt = R->dt_new;
        double wl;
        if (imex) {
            imex_weights(R->sc, R->dta, wl, h.wi, h.we);
            h.wl = wl;
        } else {
            wl = R->sc.aD[R->sc.diag * MAXS + R->sc.diag];
        }
        if (!rk && (first || R->dta[0] != R->dta[1]))
            if ((rc = rebind_ops(R, wl, R->dta[0]))) return rc;
        if ((rc = upload_desc(R))) return rc;
        if (imex) {
            imex_launches(R->dev, 1, sh, R->solo, first ? 1 : 0, R->st, nullptr);
            launch_advance(R->dev, 1, R->st);
            if (rec) launch_diag(R->dev, 1, sh, 1, nullptr, R->st);
        } else {
            step_launches(R->dev, 1, sh, R->sc, R->solo, first, restart, rec, R->st, nullptr, R->h.keep != 0);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(R->st));       // the host mirror is rewritten next step
        R->kstep++;
    }
    return 0;
}

}  // namespace

extern "C" {

int fdns_device_copy(size_t bytes, int reps, double* ms)
{
    if (bytes < 16 || reps < 1 || !ms) return fail(FDNS_EINVAL, "bytes >= 16, reps >= 1");
    const size_t n16 = bytes / 16;
    double *a = nullptr, *b = nullptr;
    if (hipMalloc((void**)&a, 16 * n16) != hipSuccess) return fail(FDNS_ENOMEM, "hipMalloc copy");
    if (hipMalloc((void**)&b, 16 * n16) != hipSuccess) { hipFree(a); return fail(FDNS_ENOMEM, "hipMalloc copy"); }
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipStreamCreate(&st);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, 16 * n16, st);
    if (e == hipSuccess) {
        launch_copy(a, b, n16, st);    // untimed first launch
        e = hipEventRecord(e0, st);
    }
    for (int k = 0; k < reps && e == hipSuccess; ++k) {
        launch_copy(a, b, n16, st);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (st) hipStreamDestroy(st);
    hipFree(a);
    hipFree(b);
    if (e != hipSuccess) return fail(FDNS_EHIP, hipGetErrorString(e));
    *ms = (double)t / reps;
    return 0;
}

const char* fdns_strerror(int code)
{
    if (g_err_lit) return g_err_lit;
    if (!g_err.empty()) return g_err.c_str();
    switch (code) {
        case FDNS_OK: return "ok";
        case FDNS_EINVAL: return "invalid parameter";
        case FDNS_ESCHEME: return "scheme not available";
        case FDNS_ESIZE: return "unsupported size";
        case FDNS_ENOMEM: return "out of memory";
        case FDNS_EHIP: return "HIP error";
        case FDNS_ESTATE: return "call out of order";
        case FDNS_ENOTFOUND: return "unknown job";
        case FDNS_EBUSY: return "pool budget exceeded";
    }
    return "unknown error";
}

size_t fdns_job_footprint(const fdns_params* p)
{
    Scheme sc;
    if (!p || scheme_lookup(p->time_scheme_type, p->time_scheme_imp, p->time_scheme_exp, sc)) return 0;
    if (p->solver < FDNS_SOLVE_AUTO || p->solver > FDNS_SOLVE_INVERSE) return 0;
    OpKey kT, kF;
    run_keys(p, sc, kT, kF);
    size_t b = run_state_bytes(p, sc);
    if (sc.kind != SK_RK) b += opset_bytes(kT.kind, p->Nr, p->Nm) + opset_bytes(kF.kind, p->Nr, p->Nm);
    if ((sc.kind == SK_IMEXRK && !sc.sa) || sc.kind == SK_RK)
        b += opset_bytes(OP_LAP, p->Nr, p->Nm) + sizeof(double) * (4 * LAGMAX + 2);
    if (sc.diag_diff) {   // the other stages' diagonals
        std::set<double> w;
        for (int k = 1; k < sc.S; ++k)
            if (k != sc.diag) w.insert(sc.aD[k * MAXS + k]);
        w.erase(sc.aD[sc.diag * MAXS + sc.diag]);
        b += w.size() * (opset_bytes(kT.kind, p->Nr, p->Nm) + opset_bytes(kF.kind, p->Nr, p->Nm));
    }
    return b;
}

int fdns_create(const fdns_params* p, fdns_run** out)
{
    if (!out) return fail(FDNS_EINVAL, "null out");
    *out = nullptr;
    g_err.clear();
    g_err_lit = nullptr;
    Scheme sc;
    int rc = validate(p, sc);
    if (rc) return rc;
    if ((rc = ensure_kernels())) return rc;
    auto* R = new fdns_run();
    R->p = *p;
    R->sc = sc;
    if (hipStreamCreateWithFlags(&R->st, hipStreamNonBlocking) != hipSuccess) { delete R; return fail(FDNS_EHIP, "stream"); }
    if ((rc = R->side.create())) { fdns_destroy(R); return rc; }
    if ((rc = shape_get(p->Nr, p->Nm, p->eta, &R->shape))) { fdns_destroy(R); return rc; }
    double wl = 0.0;
    RunDesc& h = R->h;
    std::memset(&h, 0, sizeof(h));
    const double dta[4] = {p->dt, p->dt, p->dt, p->dt};
    if (sc.kind == SK_IMEX) imex_weights(sc, dta, wl, h.wi, h.we);
    else wl = sc.aD[sc.diag * MAXS + sc.diag];
    h.wl = wl;
    OpKey kT, kF;
    double wtdt = 0.0;
    run_keys(p, sc, kT, kF, &wtdt);
    const int mBC = std::strcmp(p->mBC, "NS") ? 1 : 0;
    // the explicit RK path has no implicit operators
    if (sc.kind != SK_RK && ((rc = opset_get(R->shape, kT, wtdt, p->Pr, mBC, R->st, &R->opT)) ||
                             (rc = opset_get(R->shape, kF, wtdt, p->Pr, mBC, R->st, &R->opF)))) {
        fdns_destroy(R);
        return rc;
    }
    if ((sc.kind == SK_IMEXRK && !sc.sa) || sc.kind == SK_RK) {
        // Assembly_stage: LAPpsi of every mode (steptime_imexrk.F90:92-94,
        // steptime_rk.F90:85-87) and the Lagrange weights of init.F90:268-271
        // (nodes radius(2..lm+1) for r_min, radius(Nr_max-1..Nr_max-lm) for
        // r_max)
        const OpKey kL = op_key(OP_LAP, p->Nr, p->Nm, p->eta, 0.0, 0.0, mBC, 0);
        if ((rc = opset_get(R->shape, kL, 0.0, 0.0, mBC, R->st, &R->opL))) { fdns_destroy(R); return rc; }
        const int N = p->Nr, lm = p->lagpts;
        const double pi = 4.0 * std::atan(1.0);
        const double rmin = p->eta / (1.0 - p->eta), rmax = 1.0 / (1.0 - p->eta);
        auto rad = [&](int i) { return rmin + (0.5 * (1.0 + std::cos(((double)(N - 1 - i) * pi) / (double)(N - 1)))); };
        std::vector<double> w(4 * LAGMAX + 2, 0.0), xmin(lm), xmax(lm);
        for (int i = 0; i < lm; ++i) {
            xmin[i] = rad(1 + i);
            xmax[i] = rad(N - 2 - i);
        }
        lagrange_weights(lm, xmin.data(), rmin, w.data(), w.data() + LAGMAX);
        lagrange_weights(lm, xmax.data(), rmax, w.data() + 2 * LAGMAX, w.data() + 3 * LAGMAX);
        // stress-free walls: -2 / radius(1), -2 / radius(Nr_max) on the IMEXRK
        // path (fourierloop_imexrk.F90:634-635), -2 r_radius on the RK path
        // (fourierloop_rk.F90:206-207, 407-408)
        if (sc.kind == SK_RK) {
            w[4 * LAGMAX] = -2.0 * (1.0 / rad(0));
            w[4 * LAGMAX + 1] = -2.0 * (1.0 / rad(N - 1));
        } else {
            w[4 * LAGMAX] = -2.0 / rad(0);
            w[4 * LAGMAX + 1] = -2.0 / rad(N - 1);
        }
        if (hipMalloc((void**)&R->lagw, sizeof(double) * w.size()) != hipSuccess) { fdns_destroy(R); return fail(FDNS_ENOMEM, "hipMalloc lagrange"); }
        if (hipMemcpy(R->lagw, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice) != hipSuccess) { fdns_destroy(R); return fail(FDNS_EHIP, "lagrange upload"); }
        h.lagw = R->lagw;
        h.lm = lm;
    }
    if ((rc = stage_ops(R, p->dt))) { fdns_destroy(R); return rc; }
    h.mbc = mBC;
    h.sa = sc.sa;
    std::memcpy(h.bD, sc.bD, sizeof(h.bD));
    std::memcpy(h.bA, sc.bA, sizeof(h.bA));
    // device state
    R->bytes = run_state_bytes(p, sc);
    if (hipMalloc(&R->mem, R->bytes) != hipSuccess) { fdns_destroy(R); return fail(FDNS_ENOMEM, "hipMalloc run state"); }
    if (hipMalloc((void**)&R->own_dev, sizeof(RunDesc)) != hipSuccess) { fdns_destroy(R); return fail(FDNS_ENOMEM, "hipMalloc desc"); }
    R->dev = R->own_dev;
    const int N = p->Nr, Nm1 = p->Nm + 1;
    const size_t S = (size_t)Nm1 * N;
    char* q = (char*)R->mem;
    auto take = [&](size_t n2) { double2* r = (double2*)q; q += sizeof(double2) * n2; return r; };
    h.tFR = take(S); h.omgFR = take(S); h.upFR = take(S); h.urFR = take(S); h.psi = take(S);
    h.tspec = take(S); h.wspec = take(S);
    h.upT = take(S); h.urT = take(S); h.upW = take(S); h.urW = take(S);
    const int L = std::max(sc.n_imp, sc.n_exp);
    if (sc.kind == SK_IMEX) {
        h.hiT = take(L * S); h.heT = take(L * S); h.hiW = take(L * S); h.heW = take(L * S);
    } else if (sc.kind == SK_RK) {
        h.FdT = take(sc.S * S); h.FdW = take(sc.S * S);
        h.upC = take(S); h.d1u = take(S);
    } else {
        h.FaT = take(sc.S * S); h.FdT = take(sc.S * S); h.FaW = take(sc.S * S); h.FdW = take(sc.S * S); h.Fb = take(sc.S * S);
    }
    h.sx = take(2 * S); h.sy = take(2 * S);
    h.ubar = take(N); h.TFC = take(N); h.hiU = take((size_t)L * N); h.heU = take((size_t)L * N);
    if (p->n_KE > 0) {
        h.tpass = take(S);
        h.up0prev = take(N);
        h.keep = 1;
    }
    h.urw = (double*)q; q += sizeof(double) * N;
    h.cfl = (double*)q; q += sizeof(double) * 2 * N;
    h.solve_lu = solver_lu(p) ? 1 : 0;
    bind_ops(R);
    h.vardt = p->l_vartimestep ? 1 : 0;
    R->dt_new = p->dt;
    std::fill(R->dta, R->dta + 8, p->dt);
    h.Ra = p->Ra; h.Pr = p->Pr; h.dt = p->dt;
    h.kind = sc.kind; h.imp = sc.imp; h.n_imp = sc.n_imp; h.n_exp = sc.n_exp;
    h.L = (sc.kind == SK_IMEX) ? L : 1;
    h.head = (sc.kind == SK_IMEX) ? L - 1 : 0;
    h.buo_imp = std::strcmp(p->buo_tscheme, "IMP") == 0 ? 1 : 0;
    h.S = sc.S;
    std::memcpy(h.aD, sc.aD, sizeof(h.aD));
    std::memcpy(h.aA, sc.aA, sizeof(h.aA));
    std::memcpy(h.nl_col, sc.nl_col, sizeof(h.nl_col));
    std::memcpy(h.fa_col, sc.fa_col, sizeof(h.fa_col));
    std::memcpy(h.need_fd, sc.need_fd, sizeof(h.need_fd));
    if (p->n_KE > 0) {
        R->series_cap = 1 << 16;
        if (hipMalloc((void**)&R->series, sizeof(double) * 4 * (size_t)R->series_cap) != hipSuccess) {
            fdns_destroy(R);
            return fail(FDNS_ENOMEM, "hipMalloc series");
        }
    }
    h.series = R->series; h.cap = R->series_cap; h.nrec = 0;
    if (hipMalloc((void**)&R->d_out6, sizeof(double) * 6) != hipSuccess) { fdns_destroy(R); return fail(FDNS_ENOMEM, "hipMalloc out"); }
    if ((rc = run_zero_dynamic(R)) || (rc = upload_desc(R))) { fdns_destroy(R); return rc; }
    if ((rc = plan_build(std::vector<fdns_run*>{R}, R->solo))) { fdns_destroy(R); return rc; }
    if (hipStreamSynchronize(R->st) != hipSuccess) { fdns_destroy(R); return fail(FDNS_EHIP, "factorisation failed"); }
    *out = R;
    return 0;
}

int fdns_destroy(fdns_run* R)
{
    if (!R) return 0;
    if (R->st) hipStreamSynchronize(R->st);
    if (R->gexec) hipGraphExecDestroy(R->gexec);
    if (R->graph) hipGraphDestroy(R->graph);
    if (R->opT) opset_put(R->opT);
    if (R->opF) opset_put(R->opF);
    if (R->opL) opset_put(R->opL);
    for (int k = 0; k < MAXS; ++k) {
        if (R->stT[k]) opset_put(R->stT[k]);
        if (R->stF[k]) opset_put(R->stF[k]);
    }
    if (R->lagw) hipFree(R->lagw);
    if (R->shape) shape_put(R->shape);
    if (R->mem) hipFree(R->mem);
    if (R->own_dev) hipFree(R->own_dev);
    if (R->series) hipFree(R->series);
    if (R->d_out6) hipFree(R->d_out6);
    R->solo.release();
    R->side.destroy();
    if (R->st) hipStreamDestroy(R->st);
    delete R;
    return 0;
}

int fdns_set_state(fdns_run* R, const double* tFR, const double* omgFR, const double* urFR, const double* upFR)
{
    if (!R || !tFR || !omgFR || !urFR || !upFR) return fail(FDNS_EINVAL, "null argument");
    if (R->in_ens) return fail(FDNS_ESTATE, "run is part of an ensemble");
    const size_t S = (size_t)(R->p.Nm + 1) * R->p.Nr;
    int rc;
    if ((rc = run_zero_dynamic(R))) return rc;
    RunDesc& h = R->h;
    HIPCHK(hipMemcpyAsync(h.tFR, tFR, sizeof(double2) * S, hipMemcpyHostToDevice, R->st));
    HIPCHK(hipMemcpyAsync(h.omgFR, omgFR, sizeof(double2) * S, hipMemcpyHostToDevice, R->st));
    HIPCHK(hipMemcpyAsync(h.urFR, urFR, sizeof(double2) * S, hipMemcpyHostToDevice, R->st));
    HIPCHK(hipMemcpyAsync(h.upFR, upFR, sizeof(double2) * S, hipMemcpyHostToDevice, R->st));
    if (R->sc.kind != SK_IMEX) {
        // init_perturbation / init_fields: temp_spec = tFR, omg_spec = omgFR,
        // uphi_bar_spec = upFR(1,:) (init.F90:167-174, 393-397)
        HIPCHK(hipMemcpyAsync(h.tspec, h.tFR, sizeof(double2) * S, hipMemcpyDeviceToDevice, R->st));
        HIPCHK(hipMemcpyAsync(h.wspec, h.omgFR, sizeof(double2) * S, hipMemcpyDeviceToDevice, R->st));
        HIPCHK(hipMemcpy2DAsync(h.ubar, sizeof(double2), h.upFR, sizeof(double2) * (R->p.Nm + 1), sizeof(double2),
                                R->p.Nr, hipMemcpyDeviceToDevice, R->st));
    }
    if (h.keep) {
        // init: tt is the initial temperature, upFR_prev the initial u_phi (init.F90:253)
        HIPCHK(hipMemcpyAsync(h.tpass, h.tFR, sizeof(double2) * S, hipMemcpyDeviceToDevice, R->st));
        HIPCHK(hipMemcpy2DAsync(h.up0prev, sizeof(double2), h.upFR, sizeof(double2) * (R->p.Nm + 1), sizeof(double2),
                                R->p.Nr, hipMemcpyDeviceToDevice, R->st));
    }
    h.head = (R->sc.kind == SK_IMEX) ? h.L - 1 : 0;
    h.kstep = 0; h.time = 0.0; h.nrec = 0;
    h.hist_ready = 0;
    if (h.vardt) {
        // back to the first step's dt and weights; the operators follow at
        // the first step
        R->dt_new = R->p.dt;
        std::fill(R->dta, R->dta + 8, R->p.dt);
        h.dt = R->p.dt;
        if (R->sc.kind == SK_IMEX) imex_weights(R->sc, R->dta, h.wl, h.wi, h.we);
    }
    R->dta_restored = false;
    if ((rc = upload_desc(R))) return rc;
    HIPCHK(hipStreamSynchronize(R->st));
    R->kstep = 0;
    R->has_state = true;
    if (R->gexec) { hipGraphExecDestroy(R->gexec); R->gexec = nullptr; }
    if (R->graph) { hipGraphDestroy(R->graph); R->graph = nullptr; }
    return 0;
}

// init_fields + init_perturbation (init.F90:164-399) of a fresh start: the
// conductive temperature with the perturbation n_init selects, built in
// physical space on the (Np, Nr) grid as the reference builds tt, then
// forward-transformed level by level (forfft) on the GPU; u = omega = 0.
int fdns_init_perturbation(fdns_run* R, int n_init, double ampT)
{
    if (!R) return fail(FDNS_EINVAL, "null run");
    if (R->in_ens) return fail(FDNS_ESTATE, "run is part of an ensemble");
    const int N = R->p.Nr, Nm = R->p.Nm, Np = 3 * Nm, Nm1 = Nm + 1;
    const size_t S = (size_t)Nm1 * N;
    std::vector<double> tt;
    perturbation_field(N, Nm, R->p.eta, n_init, ampT, tt);
    double* d = nullptr;
    if (hipMalloc(&d, sizeof(double) * ((size_t)Np * N + 2 * S)) != hipSuccess)
        return fail(FDNS_ENOMEM, "hipMalloc perturbation");
    double2* fd = (double2*)(d + (size_t)Np * N);
    std::vector<double> t(2 * S), z(2 * S, 0.0);
    hipError_t e = hipMemcpyAsync(d, tt.data(), sizeof(double) * tt.size(), hipMemcpyHostToDevice, R->st);
    if (e == hipSuccess) {
        launch_forfft(d, fd, R->shape->sh, R->st);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(t.data(), fd, sizeof(double2) * S, hipMemcpyDeviceToHost, R->st);
    if (e == hipSuccess) e = hipStreamSynchronize(R->st);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(FDNS_EHIP, std::string("init_perturbation: ") + hipGetErrorString(e));
    return fdns_set_state(R, t.data(), z.data(), z.data(), z.data());
}

// Seeded initial state (build-defined; the reference draws no random numbers).
int fdns_seed_state(fdns_run* R, uint64_t seed, double amp)
{
    if (!R) return fail(FDNS_EINVAL, "null run");
    const int N = R->p.Nr, Nm = R->p.Nm, Nm1 = Nm + 1;
    const size_t S = (size_t)Nm1 * N;
    std::vector<double> t(2 * S, 0.0), z(2 * S, 0.0);
    const double eta = R->p.eta;
    const double rmin = eta / (1.0 - eta), rmax = 1.0 / (1.0 - eta);
    const double pi = 4.0 * std::atan(1.0);
    std::vector<double> rad(N);
    for (int i = 0; i < N; ++i) rad[i] = rmin + (0.5 * (1.0 + std::cos(((double)(N - 1 - i) * pi) / (double)(N - 1))));
    for (int r = 0; r < N; ++r)   // conductive profile, init.F90:207-211
        t[2 * ((size_t)Nm1 * r)] = std::log(rad[r]) / std::log(rmin / rmax) - std::log(rmax) / std::log(rmin / rmax);
    auto mix = [](uint64_t& s) {
        uint64_t x = (s += 0x9E3779B97F4A7C15ULL);
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
        return x ^ (x >> 31);
    };
    for (int m = 1; m <= Nm; ++m) {
        uint64_t s = seed * 0xD1B54A32D192ED03ULL + (uint64_t)m;
        (void)mix(s);
        const double a = (double)(mix(s) >> 11) * 0x1.0p-53 * 2.0 - 1.0;
        const double b = (double)(mix(s) >> 11) * 0x1.0p-53 * 2.0 - 1.0;
        for (int r = 1; r < N - 1; ++r) {
            const double env = std::sin(pi * (rad[r] - rmin));
            t[2 * ((size_t)m + (size_t)Nm1 * r)] = amp * a * env;
            t[2 * ((size_t)m + (size_t)Nm1 * r) + 1] = amp * b * env;
        }
    }
    return fdns_set_state(R, t.data(), z.data(), z.data(), z.data());
}

int fdns_get_state(fdns_run* R, double* tFR, double* omgFR, double* urFR, double* upFR, double* psi)
{
    if (!R) return fail(FDNS_EINVAL, "null run");
    const size_t S = (size_t)(R->p.Nm + 1) * R->p.Nr;
    const RunDesc& h = R->h;
    HIPCHK(hipStreamSynchronize(R->st));
    if (tFR) HIPCHK(hipMemcpy(tFR, h.tFR, sizeof(double2) * S, hipMemcpyDeviceToHost));
    if (omgFR) HIPCHK(hipMemcpy(omgFR, h.omgFR, sizeof(double2) * S, hipMemcpyDeviceToHost));
    if (urFR) HIPCHK(hipMemcpy(urFR, h.urFR, sizeof(double2) * S, hipMemcpyDeviceToHost));
    if (upFR) HIPCHK(hipMemcpy(upFR, h.upFR, sizeof(double2) * S, hipMemcpyDeviceToHost));
    if (psi) HIPCHK(hipMemcpy(psi, h.psi, sizeof(double2) * S, hipMemcpyDeviceToHost));
    return 0;
}

int fdns_step(fdns_run* R, int nsteps)
{
    if (!R) return fail(FDNS_EINVAL, "null run");
    if (!R->has_state) return fail(FDNS_ESTATE, "fdns_set_state first");
    if (R->in_ens) return fail(FDNS_ESTATE, "run is part of an ensemble");
    if (nsteps <= 0) return 0;
    if (R->h.vardt) return vardt_steps(R, nsteps);
    const Shape& sh = R->shape->sh;
    const bool restart = R->p.l_restart != 0;
    const int nKE = R->p.n_KE;
    int done = 0;
    if (R->kstep == 0) {
        prologue_launches(R->dev, 1, sh, R->sc, restart, R->st);
        const bool rec = nKE > 0 && (1 % nKE) == 0;
        step_launches(R->dev, 1, sh, R->sc, R->solo, true, restart, rec, R->st, nullptr, R->h.keep != 0, &R->side);
        R->kstep = 1;
        done = 1;
    }
    // steady-state steps: replay one captured step when no record is due
    while (done < nsteps) {
        const bool rec = nKE > 0 && ((R->kstep + 1) % nKE) == 0;
        if (!rec) {
            if (R->gexec && R->tune_epoch != g_tune_epoch) {
                hipGraphExecDestroy(R->gexec);
                hipGraphDestroy(R->graph);
                R->gexec = nullptr;
                R->graph = nullptr;
            }
            if (!R->gexec) {
                R->tune_epoch = g_tune_epoch;
                HIPCHK(hipStreamBeginCapture(R->st, hipStreamCaptureModeThreadLocal));
                step_launches(R->dev, 1, sh, R->sc, R->solo, false, restart, false, R->st, nullptr, R->h.keep != 0,
                              &R->side);
                HIPCHK(hipStreamEndCapture(R->st, &R->graph));
                HIPCHK(hipGraphInstantiate(&R->gexec, R->graph, nullptr, nullptr, 0));
            }
            HIPCHK(hipGraphLaunch(R->gexec, R->st));
        } else {
            step_launches(R->dev, 1, sh, R->sc, R->solo, false, restart, true, R->st, nullptr, R->h.keep != 0,
                          &R->side);
        }
        R->kstep++;
        done++;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(R->st));
    return 0;
}

int fdns_get_dt(fdns_run* R, double* dt_new, double* dt_array)
{
    if (!R || !dt_new) return fail(FDNS_EINVAL, "null argument");
    *dt_new = R->dt_new;
    if (dt_array) std::copy(R->dta, R->dta + std::max(R->sc.n_imp, R->sc.n_exp), dt_array);
    return 0;
}

// writeKE_spectral's further outputs (output.F90:122-324) of the state after
// the last step: k_budget_dr, k_budget on the run's stream.
int fdns_budget(fdns_run* R, double* out4, double* upbal)
{
    if (!R || !out4) return fail(FDNS_EINVAL, "null argument");
    if (R->in_ens) return fail(FDNS_ESTATE, "run is part of an ensemble");
    if (!R->h.keep) return fail(FDNS_ESTATE, "the budget needs a run created with n_KE > 0");
    if (R->kstep == 0) return fail(FDNS_ESTATE, "the budget needs a completed step");
    const int N = R->p.Nr;
    double* d = nullptr;
    const size_t n = 4 + 3 * (size_t)N;
    if (hipMalloc((void**)&d, sizeof(double) * n) != hipSuccess) return fail(FDNS_ENOMEM, "hipMalloc budget");
    std::vector<double> h(n);
    launch_budget(R->dev, 1, R->shape->sh, d, R->st);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d, sizeof(double) * n, hipMemcpyDeviceToHost, R->st);
    if (e == hipSuccess) e = hipStreamSynchronize(R->st);
    hipFree(d);
    if (e != hipSuccess) return fail(FDNS_EHIP, hipGetErrorString(e));
    std::copy(h.begin(), h.begin() + 4, out4);
    if (upbal) std::copy(h.begin() + 4, h.end(), upbal);
    return 0;
}

// fdns_budget without any allocation: the outputs are staged in the library's
// device scratch array (one per GPU, shared by every call on that GPU) and
// copied straight into the caller's arrays.  k_budget assigns every staged
// value, so nothing needs clearing between calls; calls that overlap on one
// GPU race on the array.
int fdns_budget_noalloc(fdns_run* R, double* out4, double* upbal)
{
    if (!R || !out4) return fail_lit(FDNS_EINVAL, "null argument");
    if (R->in_ens) return fail_lit(FDNS_ESTATE, "run is part of an ensemble");
    if (!R->h.keep) return fail_lit(FDNS_ESTATE, "the budget needs a run created with n_KE > 0");
    if (R->kstep == 0) return fail_lit(FDNS_ESTATE, "the budget needs a completed step");
    const int N = R->p.Nr;
    const size_t n = 4 + 3 * (size_t)N;
    double* d = nullptr;
    size_t cap = 0;
    if (budget_scratch(&d, &cap)) return fail_lit(FDNS_EHIP, "hipGetSymbolAddress of the budget scratch array failed");
    if (n > cap) return fail_lit(FDNS_ESIZE, "the budget scratch array holds 4 + 3 * 513 doubles (Nr_max <= 513)");
    launch_budget(R->dev, 1, R->shape->sh, d, R->st);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out4, d, sizeof(double) * 4, hipMemcpyDeviceToHost, R->st);
    if (e == hipSuccess && upbal)
        e = hipMemcpyAsync(upbal, d + 4, sizeof(double) * 3 * (size_t)N, hipMemcpyDeviceToHost, R->st);
    if (e == hipSuccess) e = hipStreamSynchronize(R->st);
    if (e != hipSuccess) return fail_lit(FDNS_EHIP, hipGetErrorString(e));
    return 0;
}

// calculate_spectra (output.F90:327-355) of the current state.
int fdns_spectra(fdns_run* R, double* KEspec, double* Tspec)
{
    if (!R || !KEspec || !Tspec) return fail(FDNS_EINVAL, "null argument");
    if (R->in_ens) return fail(FDNS_ESTATE, "run is part of an ensemble");
    if (!R->has_state) return fail(FDNS_ESTATE, "fdns_set_state first");
    const int Nm1 = R->p.Nm + 1;
    double* d = nullptr;
    if (hipMalloc((void**)&d, sizeof(double) * 2 * Nm1) != hipSuccess) return fail(FDNS_ENOMEM, "hipMalloc spectra");
    std::vector<double> h(2 * (size_t)Nm1);
    launch_spectra(R->dev, 1, R->shape->sh, d, R->st);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d, sizeof(double) * 2 * Nm1, hipMemcpyDeviceToHost, R->st);
    if (e == hipSuccess) e = hipStreamSynchronize(R->st);
    hipFree(d);
    if (e != hipSuccess) return fail(FDNS_EHIP, hipGetErrorString(e));
    std::copy(h.begin(), h.begin() + Nm1, KEspec);
    std::copy(h.begin() + Nm1, h.end(), Tspec);
    return 0;
}

int fdns_diag(fdns_run* R, double* out6)
{
    if (!R || !out6) return fail(FDNS_EINVAL, "null argument");
    launch_diag(R->dev, 1, R->shape->sh, 0, R->d_out6, R->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out6, R->d_out6, sizeof(double) * 6, hipMemcpyDeviceToHost, R->st));
    HIPCHK(hipStreamSynchronize(R->st));
    return 0;
}

int fdns_get_series(fdns_run* R, double* out, int cap_rows, int* nrows)
{
    if (!R || !nrows) return fail(FDNS_EINVAL, "null argument");
    int rc;
    if ((rc = download_desc(R))) return rc;
    const int n = std::min(R->h.nrec, cap_rows);
    if (out && n > 0) HIPCHK(hipMemcpy(out, R->series, sizeof(double) * 4 * (size_t)n, hipMemcpyDeviceToHost));
    *nrows = R->h.nrec;
    return 0;
}

// ---- restart and checkpoint / snapshot files
int fdns_scheme_orders(const fdns_params* p, int* n_imp, int* n_exp)
{
    if (!p || !n_imp || !n_exp) return fail(FDNS_EINVAL, "null argument");
    Scheme sc;
    if (scheme_lookup(p->time_scheme_type, p->time_scheme_imp, p->time_scheme_exp, sc))
        return fail(FDNS_ESCHEME, "scheme not available");
    *n_imp = sc.n_imp;
    *n_exp = sc.n_exp;
    return 0;
}

int fdns_restart(fdns_run* R, int nstates, const double* tFR, const double* omgFR, const double* urFR,
                 const double* upFR, double time)
{
    if (!R || !tFR || !omgFR || !urFR || !upFR) return fail(FDNS_EINVAL, "null argument");
    if (R->in_ens) return fail(FDNS_ESTATE, "run is part of an ensemble");
    if (!R->p.l_restart) return fail(FDNS_ESTATE, "fdns_restart needs a run created with l_restart = 1");
    const bool imex = R->sc.kind == SK_IMEX;
    const int K = nstates, ni = R->sc.n_imp, ne = R->sc.n_exp;
    if (imex ? (K != 1 && K != ni) : K != 1)
        return fail(FDNS_EINVAL, "restart takes 1 checkpoint state, or n_order_tscheme_imp of them for IMEX");
    const size_t S = (size_t)(R->p.Nm + 1) * R->p.Nr;
    const size_t last = 2 * S * (size_t)(K - 1);
    int rc;
    // newest state current, histories zero (fdns_set_state)
    if ((rc = fdns_set_state(R, tFR + last, omgFR + last, urFR + last, upFR + last))) return rc;
    RunDesc& h = R->h;
    if (imex && K > 1) {
        // steptime.F90:88-186: the j-th state (oldest first) through a nonlinear
        // pass into implicit slot n_imp - j and explicit slot n_exp - j; the
        // last state is the newest, so the fields end holding it
        const Shape& sh = R->shape->sh;
        for (int j = 0; j < K; ++j) {
            const size_t o = 2 * S * (size_t)j;
            HIPCHK(hipMemcpyAsync(h.tFR, tFR + o, sizeof(double2) * S, hipMemcpyHostToDevice, R->st));
            HIPCHK(hipMemcpyAsync(h.omgFR, omgFR + o, sizeof(double2) * S, hipMemcpyHostToDevice, R->st));
            HIPCHK(hipMemcpyAsync(h.urFR, urFR + o, sizeof(double2) * S, hipMemcpyHostToDevice, R->st));
            HIPCHK(hipMemcpyAsync(h.upFR, upFR + o, sizeof(double2) * S, hipMemcpyHostToDevice, R->st));
            launch_nonlin(R->dev, 1, sh, R->st);
            launch_imex_restart(R->dev, 1, sh, R->st, ni - 1 - j, ne - 1 - j);
        }
        HIPCHK(hipGetLastError());
    }
    h.hist_ready = imex ? 1 : 0;
    h.time = time;
    if ((rc = upload_desc(R))) return rc;
    HIPCHK(hipStreamSynchronize(R->st));
    return 0;
}

int fdns_restart_dt(fdns_run* R, int n, const double* dt_array)
{
    if (!R || !dt_array) return fail(FDNS_EINVAL, "null argument");
    if (!R->p.l_restart || !R->has_state || R->kstep != 0)
        return fail(FDNS_ESTATE, "fdns_restart_dt follows fdns_restart, before the first step");
    const int nmax = std::max(R->sc.n_imp, R->sc.n_exp);
    if (n != nmax) return fail(FDNS_EINVAL, "dt_array needs max(n_order_tscheme_imp, n_order_tscheme_exp) entries");
    for (int i = 0; i < n; ++i)
        if (!(dt_array[i] > 0.0)) return fail(FDNS_EINVAL, "dt_array entries must be positive");
    std::copy(dt_array, dt_array + n, R->dta);
    R->dta_restored = true;
    return 0;
}

int fdns_get_history(fdns_run* R, double* imp_temp, double* exp_temp, double* imp_vort, double* exp_vort,
                     double* imp_ub, double* exp_ub)
{
    if (!R || !imp_temp || !exp_temp || !imp_vort || !exp_vort || !imp_ub || !exp_ub)
        return fail(FDNS_EINVAL, "null argument");
    const int N = R->p.Nr, Nm1 = R->p.Nm + 1;
    const size_t S = (size_t)Nm1 * N;
    if (R->sc.kind == SK_IMEXRK) {
        // the (1, Nm+1, Nr) arrays of the IMEXRK path, never assigned there
        for (double* a : {imp_temp, exp_temp, imp_vort, exp_vort}) std::memset(a, 0, sizeof(double2) * S);
        for (double* a : {imp_ub, exp_ub}) std::memset(a, 0, sizeof(double2) * N);
        return 0;
    }
    if (R->sc.kind == SK_RK) {
        // n_order_tscheme slots as store_checkpoint writes them; zero here.
        // The reference's rhs_imp_* are one-slot arrays written past their
        // end there, and only the first slot of rhs_exp_* is ever assigned
        // (dt times the last stage's weighted sum, fourierloop_rk.F90:120,
        // 147, 213); restarts never read them.
        const size_t n = (size_t)R->sc.S;
        for (double* a : {imp_temp, exp_temp, imp_vort, exp_vort}) std::memset(a, 0, sizeof(double2) * S * n);
        for (double* a : {imp_ub, exp_ub}) std::memset(a, 0, sizeof(double2) * N * n);
        return 0;
    }
    int rc;
    if ((rc = download_desc(R))) return rc;
    const RunDesc& h = R->h;
    const int L = h.L, ni = R->sc.n_imp, ne = R->sc.n_exp;
    // reference slot k (0-based, k = 0 newest) lives in ring slot head - k; before
    // the first step the slot the first step writes (ring slot 0) counts as
    // the newest.  The m = 0 vorticity rows are written only by the restart
    // start-up and never shifted (rhs_construct_vort runs for m >= 1,
    // fourierloop.F90:145-171), so they keep the start-up's placement.
    const int head = h.kstep == 0 ? 0 : h.head;
    auto ring = [&](int hd, int k) { return (size_t)((hd - k + 2 * L) % L); };
    std::vector<double2> buf((size_t)L * S);
    auto spec = [&](GPtr<double2> src, int n, bool vort, double* out) -> int {
        HIPCHK(hipMemcpy(buf.data(), src.p, sizeof(double2) * L * S, hipMemcpyDeviceToHost));
        double2* o = (double2*)out;
        for (size_t e = 0; e < S; ++e) {
            const int hd = (vort && e % Nm1 == 0) ? 0 : head;
            for (int k = 0; k < n; ++k) o[(size_t)k + (size_t)n * e] = buf[ring(hd, k) * S + e];
        }
        return 0;
    };
    auto mean = [&](GPtr<double2> src, int n, double* out) -> int {
        HIPCHK(hipMemcpy(buf.data(), src.p, sizeof(double2) * L * N, hipMemcpyDeviceToHost));
        double2* o = (double2*)out;
        for (int i = 0; i < N; ++i)
            for (int k = 0; k < n; ++k) o[(size_t)k + (size_t)n * i] = buf[ring(head, k) * N + i];
        return 0;
    };
    if ((rc = spec(h.hiT, ni, false, imp_temp)) || (rc = spec(h.heT, ne, false, exp_temp)) ||
        (rc = spec(h.hiW, ni, true, imp_vort)) || (rc = spec(h.heW, ne, true, exp_vort)) ||
        (rc = mean(h.hiU, ni, imp_ub)) || (rc = mean(h.heU, ne, exp_ub)))
        return rc;
    return 0;
}

}  // extern "C"

namespace {

// Fortran unformatted sequential records as gfortran and flang write them: the
// record's byte length as a little-endian int32 before and after its bytes.
struct FortranWriter {
    FILE* f = nullptr;
    bool ok = true;
    explicit FortranWriter(const char* path) { f = std::fopen(path, "wb"); ok = f != nullptr; }
    ~FortranWriter() { if (f) std::fclose(f); }
    void put(const void* p, size_t n) { if (ok && n && std::fwrite(p, 1, n, f) != n) ok = false; }
    // one record from several items
    void record(std::initializer_list<std::pair<const void*, size_t>> items)
    {
        size_t n = 0;
        for (auto& it : items) n += it.second;
        if (n > 0x7fffffffu) { ok = false; return; }   // would need the split-record form
        const int32_t m = (int32_t)n;
        put(&m, 4);
        for (auto& it : items) put(it.first, it.second);
        put(&m, 4);
    }
    int close()
    {
        if (f && std::fclose(f) != 0) ok = false;
        f = nullptr;
        return ok ? 0 : fail(FDNS_EINVAL, "checkpoint/snapshot file write failed");
    }
};

}  // namespace

extern "C" {

// Restart with a change of resolution (init.F90:722-861): the four fields of
// an Nm_old x Nr_old checkpoint on Nm_new x Nr_new.  Per mode, FFTW's
// REDFT00 (chebtransform_redft, chebyshev_fftw.F90:196-216) on the old grid,
// zero-padded, REDFT00 on the new grid over 2 (Nr_new - 1)
// (chebinvtran_redft, :218-245), times 2 (Nr_new - 1) / (2 (Nr_old - 1));
// modes above Nm_old zero.  The transforms are summed directly (a one-time
// conversion).  As there, the highest old Chebyshev mode enters twice its
// weight; coefficients the reference leaves unassigned are zero here.
int fdns_refine_state(int Nm_old, int Nr_old, const double* tFR, const double* omgFR, const double* urFR,
                      const double* upFR, int Nm_new, int Nr_new, double* tN, double* omgN, double* urN, double* upN)
{
    if (!tFR || !omgFR || !urFR || !upFR || !tN || !omgN || !urN || !upN) return fail(FDNS_EINVAL, "null argument");
    if (Nm_old < 1 || Nr_old < 4 || Nm_new < Nm_old || Nr_new < Nr_old)
        return fail(FDNS_ESIZE, "a resolution change refines (Nm_new >= Nm_old, Nr_new >= Nr_old)");
    const double* in[4] = {tFR, omgFR, urFR, upFR};
    double* out[4] = {tN, omgN, urN, upN};
    const size_t So = (size_t)(Nm_old + 1) * Nr_old, Sn = (size_t)(Nm_new + 1) * Nr_new;
    if (Nm_new == Nm_old && Nr_new == Nr_old) {
        for (int f = 0; f < 4; ++f) std::copy(in[f], in[f] + 2 * So, out[f]);
        return 0;
    }
    const double pi = 4.0 * std::atan(1.0);
    auto redft00 = [&](int n, const double* x, double* y) {
        const long L = 2 * (long)(n - 1);
        for (int k = 0; k < n; ++k) {
            double v = x[0] + ((k % 2) ? -x[n - 1] : x[n - 1]);
            for (int j = 1; j < n - 1; ++j) v = v + 2.0 * x[j] * std::cos(pi * (double)(((long)j * k) % L) / (double)(n - 1));
            y[k] = v;
        }
    };
    const double fac = (2.0 * (double)(Nr_new - 1)) / (2.0 * (double)(Nr_old - 1));
    std::vector<double> xo(Nr_old), co(Nr_old), cn(Nr_new), yn(Nr_new);
    for (int f = 0; f < 4; ++f) {
        std::fill(out[f], out[f] + 2 * Sn, 0.0);
        for (int m = 0; m <= Nm_old; ++m)
            for (int part = 0; part < 2; ++part) {
                for (int r = 0; r < Nr_old; ++r) xo[r] = in[f][2 * (m + (size_t)(Nm_old + 1) * r) + part];
                redft00(Nr_old, xo.data(), co.data());
                for (int r = 0; r < Nr_new; ++r) cn[r] = r < Nr_old ? co[r] : 0.0;
                redft00(Nr_new, cn.data(), yn.data());
                for (int r = 0; r < Nr_new; ++r)
                    out[f][2 * (m + (size_t)(Nm_new + 1) * r) + part] = (yn[r] / (2.0 * (double)(Nr_new - 1))) * fac;
            }
    }
    return 0;
}

int fdns_write_checkpoint(fdns_run* R, const char* path)
{
    if (!R || !path) return fail(FDNS_EINVAL, "null argument");
    if (!R->has_state) return fail(FDNS_ESTATE, "fdns_set_state first");
    const int N = R->p.Nr, Nm = R->p.Nm;
    const size_t S = (size_t)(Nm + 1) * N;
    const bool imex = R->sc.kind == SK_IMEX;
    const int32_t ni = R->sc.n_imp, ne = R->sc.n_exp, nmax = std::max(ni, ne);
    // slots of the stored history arrays: the IMEX orders; one on the IMEXRK
    // path (store_checkpoint_imexrk); the stage count on the RK path, whose
    // store_checkpoint declares them (n_order_tscheme_imp/exp, Nm+1, Nr)
    const bool rk = R->sc.kind == SK_RK;
    const int hi = (imex || rk) ? ni : 1, he = (imex || rk) ? ne : 1;
    std::vector<double> st(8 * S), hist(2 * S * (size_t)(2 * hi + 2 * he) + 2 * (size_t)N * (hi + he));
    int rc;
    if ((rc = download_desc(R))) return rc;
    if ((rc = fdns_get_state(R, st.data(), st.data() + 2 * S, st.data() + 4 * S, st.data() + 6 * S, nullptr))) return rc;
    double* it = hist.data();
    double* et = it + 2 * S * hi;
    double* iw = et + 2 * S * he;
    double* ew = iw + 2 * S * hi;
    double* iu = ew + 2 * S * he;
    double* eu = iu + 2 * (size_t)N * hi;
    if ((rc = fdns_get_history(R, it, et, iw, ew, iu, eu))) return rc;
    const double dt = R->dt_new, t = R->h.time;
    const int32_t nm = Nm, nr = N;
    std::vector<double> dta(R->dta, R->dta + nmax);
    const size_t F = sizeof(double2) * S;
    // store_checkpoint / store_checkpoint_imexrk (output.F90:359-450)
    FortranWriter w(path);
    if (!w.ok) return fail(FDNS_EINVAL, std::string("cannot open ") + path);
    w.record({{&dt, 8}, {&t, 8}, {&nm, 4}, {&nr, 4}, {&ni, 4}, {&ne, 4}, {&nmax, 4}});
    w.record({{dta.data(), 8 * (size_t)nmax}});
    for (int f = 0; f < 4; ++f) w.record({{st.data() + 2 * S * f, F}});
    w.record({{it, F * hi}});
    w.record({{et, F * he}});
    w.record({{iw, F * hi}});
    w.record({{ew, F * he}});
    w.record({{iu, sizeof(double2) * (size_t)N * hi}});
    w.record({{eu, sizeof(double2) * (size_t)N * he}});
    return w.close();
}

int fdns_write_snapshot(fdns_run* R, const char* path)
{
    if (!R || !path) return fail(FDNS_EINVAL, "null argument");
    if (!R->has_state) return fail(FDNS_ESTATE, "fdns_set_state first");
    const int N = R->p.Nr, Nm = R->p.Nm;
    const size_t S = (size_t)(Nm + 1) * N;
    std::vector<double> st(8 * S);
    int rc;
    if ((rc = download_desc(R))) return rc;
    if ((rc = fdns_get_state(R, st.data(), st.data() + 2 * S, st.data() + 4 * S, st.data() + 6 * S, nullptr))) return rc;
    // the grid of init_grid (init.F90:129-131)
    const double pi = 4.0 * std::atan(1.0);
    const double rmin = R->p.eta / (1.0 - R->p.eta);
    std::vector<double> rad(N);
    for (int i = 0; i < N; ++i) rad[i] = rmin + (0.5 * (1.0 + std::cos(((double)(N - 1 - i) * pi) / (double)(N - 1))));
    const double dt = R->dt_new, t = R->h.time;
    const int32_t nm = Nm, nr = N;
    const size_t F = sizeof(double2) * S;
    // store_snapshot (output.F90:488-516); the IMEXRK variant adds omgFR_check,
    // which that path never assigns (output.F90:451-485; fourierloop_imexrk.F90:377)
    FortranWriter w(path);
    if (!w.ok) return fail(FDNS_EINVAL, std::string("cannot open ") + path);
    w.record({{&R->p.Ra, 8}, {&R->p.Pr, 8}, {&R->p.eta, 8}});
    w.record({{&dt, 8}, {&t, 8}});
    w.record({{&nm, 4}, {&nr, 4}});
    w.record({{rad.data(), 8 * (size_t)N}});
    for (int f = 0; f < 4; ++f) w.record({{st.data() + 2 * S * f, F}});
    if (R->sc.kind == SK_IMEXRK) {
        std::vector<double> z(2 * S, 0.0);
        w.record({{z.data(), F}});
    }
    return w.close();
}

}  // extern "C"

// ------------------------------------------------------------------ ensembles
namespace {

struct Group {
    ShapeCtx* shape;
    Scheme sc;          // scheme structure shared by the group
    bool restart;
    std::vector<fdns_run*> runs;
    RunDesc* dev = nullptr;   // contiguous descriptors of the group
    SolveSched solve;         // batched solves of the group
    bool keep = false;        // a run of the group keeps the budget inputs (n_KE > 0)
};

bool same_structure(const fdns_run* a, const fdns_run* b)
{
    if (a->shape != b->shape) return false;
    if (a->sc.kind != b->sc.kind) return false;
    if (a->h.solve_lu != b->h.solve_lu) return false;
    if (a->kstep != b->kstep && (a->kstep == 0 || b->kstep == 0)) return false;
    // a restart changes only a run's first step (its prologue and first
    // nonlinear pass): past it, restarted and fresh runs share one launch
    if (a->kstep == 0 && (a->p.l_restart != 0) != (b->p.l_restart != 0)) return false;
    if (a->sc.kind == SK_IMEX) return true;   // one IMEX kernel serves every IMEX scheme
    if (a->sc.S != b->sc.S || a->sc.sa != b->sc.sa) return false;
    for (int k = 0; k < MAXS; ++k)
        if (a->sc.nl_col[k] != b->sc.nl_col[k]) return false;
    return true;
}

}  // namespace

// Independent groups of a batched step run on up to K streams, forked from and
// joined back into the owner's stream through events (hipGraph capture records
// the same fork and join).  All streams keep the default priority.
struct Lanes {
    static constexpr int K = 4;
    hipStream_t s[K] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[K] = {};
    Side side[K];
    int create()
    {
        for (int k = 0; k < K; ++k) {
            if (hipStreamCreateWithFlags(&s[k], hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&join[k], hipEventDisableTiming) != hipSuccess)
                return fail(FDNS_EHIP, "lane streams");
            if (int rc = side[k].create()) return rc;
        }
        if (hipEventCreateWithFlags(&fork, hipEventDisableTiming) != hipSuccess) return fail(FDNS_EHIP, "lane events");
        return 0;
    }
    void destroy()
    {
        for (int k = 0; k < K; ++k) {
            side[k].destroy();
            if (s[k]) hipStreamDestroy(s[k]);
            if (join[k]) hipEventDestroy(join[k]);
            s[k] = nullptr;
            join[k] = nullptr;
        }
        if (fork) hipEventDestroy(fork);
        fork = nullptr;
    }
};

// One batched step of every group: group i on lane i mod K when there are
// several groups.  A group takes its first step (with the prologue) when its
// runs have taken none.
// With a timer every group runs on the main stream, one after another, so that
// each event pair brackets its kernel alone on the chip.
void groups_step(std::vector<Group>& groups, hipStream_t main, Lanes& L, bool rec, KTimer* kt = nullptr)
{
    const int used = kt ? 1 : std::min<int>(Lanes::K, (int)groups.size());
    if (used > 1) {
        hipEventRecord(L.fork, main);
        for (int k = 0; k < used; ++k) hipStreamWaitEvent(L.s[k], L.fork, 0);
    }
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        Group& g = groups[gi];
        hipStream_t st = used > 1 ? L.s[gi % used] : main;
        const int n = (int)g.runs.size();
        const bool first = g.runs[0]->kstep == 0;
        if (first) prologue_launches(g.dev, n, g.shape->sh, g.sc, g.restart, st);
        step_launches(g.dev, n, g.shape->sh, g.sc, g.solve, first, g.restart, rec, st, kt, g.keep,
                      kt ? nullptr : &L.side[used > 1 ? gi % used : 0]);
    }
    if (used > 1) {
        for (int k = 0; k < used; ++k) {
            hipEventRecord(L.join[k], L.s[k]);
            hipStreamWaitEvent(main, L.join[k], 0);
        }
    }
}

// The captured steady-state batched step: one graph per group, captured on
// and replayed into the group's lane stream.  (One graph of every group is
// executed by the runtime in capture order over its own streams, which puts
// the later groups' first kernels behind the earlier groups' kernels; one
// graph per lane keeps the groups' chains side by side.)
struct StepGraphs {
    std::vector<hipGraph_t> g;
    std::vector<hipGraphExec_t> x;
    bool empty() const { return x.empty(); }
    void release()
    {
        for (auto e : x) hipGraphExecDestroy(e);
        for (auto h : g) hipGraphDestroy(h);
        x.clear();
        g.clear();
    }
    int capture(std::vector<Group>& groups, hipStream_t main, Lanes& L)
    {
        release();
        const int used = std::min<int>(Lanes::K, (int)groups.size());
        for (size_t gi = 0; gi < groups.size(); ++gi) {
            Group& grp = groups[gi];
            hipStream_t st = used > 1 ? L.s[gi % used] : main;
            hipGraph_t h = nullptr;
            hipGraphExec_t e = nullptr;
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            step_launches(grp.dev, (int)grp.runs.size(), grp.shape->sh, grp.sc, grp.solve, false, grp.restart, false, st,
                          nullptr, grp.keep, &L.side[used > 1 ? gi % used : 0]);
            HIPCHK(hipStreamEndCapture(st, &h));
            g.push_back(h);
            HIPCHK(hipGraphInstantiate(&e, h, nullptr, nullptr, 0));
            x.push_back(e);
        }
        return 0;
    }
    // reps batched steps: fork from main, each group's graph replayed reps
    // times on its lane (the groups are independent, so a lane runs ahead of
    // the others up to the join), join
    int launch(hipStream_t main, Lanes& L, int reps = 1)
    {
        const int used = std::min<int>(Lanes::K, (int)x.size());
        if (used > 1) {
            HIPCHK(hipEventRecord(L.fork, main));
            for (int k = 0; k < used; ++k) HIPCHK(hipStreamWaitEvent(L.s[k], L.fork, 0));
        }
        // interleaved, so that every lane has its next step queued while the
        // host submits the others'
        for (int r = 0; r < reps; ++r)
            for (size_t gi = 0; gi < x.size(); ++gi) HIPCHK(hipGraphLaunch(x[gi], used > 1 ? L.s[gi % used] : main));
        if (used > 1) {
            for (int k = 0; k < used; ++k) {
                HIPCHK(hipEventRecord(L.join[k], L.s[k]));
                HIPCHK(hipStreamWaitEvent(main, L.join[k], 0));
            }
        }
        return 0;
    }
};

struct fdns_ens {
    std::vector<Group> groups;
    Lanes lanes;
    hipStream_t st = nullptr;
    StepGraphs sg;        // the captured steady-state step
    int kstep = 0;
    int nKE = 0;
    int tune_epoch = 0;   // launch choices of the solve plans and the graph
    std::vector<fdns_run*> members;   // in the caller's order
    bool grouped_past_first = false;  // groups formed after the first step
};

namespace {

// Solve plans and a captured graph made before the last change of the launch
// choices are made again.
int groups_retune(std::vector<Group>& groups, int& epoch, StepGraphs& sg)
{
    const int now = g_tune_epoch;
    if (epoch == now) return 0;
    sg.release();
    for (auto& g : groups)
        if (!g.runs.empty())
            if (int rc = plan_build(g.runs, g.solve)) return rc;
    epoch = now;
    return 0;
}

// Groups of the members by structure, in the members' order.
void ens_group(fdns_ens* E)
{
    E->groups.clear();
    for (fdns_run* R : E->members) {
        bool placed = false;
        for (auto& g : E->groups)
            if (same_structure(g.runs[0], R)) { g.runs.push_back(R); g.keep = g.keep || R->h.keep; placed = true; break; }
        if (!placed) {
            Group g;
            g.shape = R->shape; g.sc = R->sc; g.restart = R->p.l_restart != 0; g.keep = R->h.keep != 0;
            g.runs.push_back(R);
            E->groups.push_back(g);
        }
    }
    E->grouped_past_first = E->members[0]->kstep > 0;
    // past the first step, with lane_split: while lanes are free, halve the
    // group with the most work per step (runs x stages) when it has more than
    // its share of the total and at least 8 runs
    if (!E->grouped_past_first) return;
    auto work = [](const Group& g) {
        return (double)g.runs.size() * (g.sc.kind == SK_IMEX ? 1.0 : (double)g.sc.S);
    };
    while ((int)E->groups.size() < Lanes::K) {
        double total = 0.0;
        int best = -1;
        for (size_t i = 0; i < E->groups.size(); ++i) {
            total += work(E->groups[i]);
            if (E->groups[i].shape->sh.lane_split && E->groups[i].runs.size() >= 8 &&
                (best < 0 || work(E->groups[i]) > work(E->groups[best])))
                best = (int)i;
        }
        if (best < 0 || work(E->groups[best]) <= total / Lanes::K) break;
        Group& g = E->groups[best];
        Group h;
        h.shape = g.shape; h.sc = g.sc; h.restart = g.restart;
        const size_t half = g.runs.size() / 2;
        h.runs.assign(g.runs.begin() + half, g.runs.end());
        g.runs.resize(half);
        g.keep = h.keep = false;
        for (auto* R : g.runs) g.keep = g.keep || R->h.keep;
        for (auto* R : h.runs) h.keep = h.keep || R->h.keep;
        E->groups.push_back(h);
    }
}

void ens_unbind(fdns_ens* E);
int ens_bind(fdns_ens* E);

// After the first step the groups that differed only in the restart flag are
// one structure: form the groups again (the captured graph, made with the old
// groups, is dropped).
int ens_regroup(fdns_ens* E)
{
    if (E->kstep == 0) return 0;
    // also when the launch choices changed (lane_split forms the groups)
    if (E->grouped_past_first && E->tune_epoch == g_tune_epoch) return 0;
    HIPCHK(hipStreamSynchronize(E->st));
    E->sg.release();
    ens_unbind(E);
    ens_group(E);
    return ens_bind(E);
}

int ens_bind(fdns_ens* E)
{
    E->tune_epoch = g_tune_epoch;
    for (auto& g : E->groups) {
        const int n = (int)g.runs.size();
        if (hipMalloc((void**)&g.dev, sizeof(RunDesc) * n) != hipSuccess) return fail(FDNS_ENOMEM, "ensemble descriptors");
        for (int i = 0; i < n; ++i) {
            fdns_run* R = g.runs[i];
            HIPCHK(hipStreamSynchronize(R->st));
            HIPCHK(hipMemcpy(g.dev + i, R->dev, sizeof(RunDesc), hipMemcpyDeviceToDevice));
            R->dev = g.dev + i;
            R->in_ens = true;
        }
        int rc = plan_build(g.runs, g.solve);
        if (rc) return rc;
    }
    return 0;
}

void ens_unbind(fdns_ens* E)
{
    for (auto& g : E->groups) {
        for (size_t i = 0; i < g.runs.size(); ++i) {
            fdns_run* R = g.runs[i];
            hipMemcpy(R->own_dev, g.dev + i, sizeof(RunDesc), hipMemcpyDeviceToDevice);
            R->dev = R->own_dev;
            R->in_ens = false;
            download_desc(R);
        }
        hipFree(g.dev);
        g.dev = nullptr;
        g.solve.release();
    }
}

}  // namespace

extern "C" {

int fdns_ens_create(fdns_run* const* runs, int nruns, fdns_ens** out)
{
    if (!runs || nruns <= 0 || !out) return fail(FDNS_EINVAL, "bad ensemble arguments");
    auto* E = new fdns_ens();
    if (hipStreamCreateWithFlags(&E->st, hipStreamNonBlocking) != hipSuccess) { delete E; return fail(FDNS_EHIP, "stream"); }
    for (int i = 0; i < nruns; ++i) {
        fdns_run* R = runs[i];
        if (!R || !R->has_state || R->in_ens) { hipStreamDestroy(E->st); delete E; return fail(FDNS_ESTATE, "ensemble member not ready"); }
        if (R->h.vardt) { hipStreamDestroy(E->st); delete E; return fail(FDNS_EINVAL, "variable-dt runs step alone"); }
        if (R->kstep != runs[0]->kstep) { hipStreamDestroy(E->st); delete E; return fail(FDNS_ESTATE, "members must be at the same step"); }
        E->members.push_back(R);
    }
    ens_group(E);
    E->kstep = runs[0]->kstep;
    E->nKE = runs[0]->p.n_KE;
    int rc = ens_bind(E);
    if (!rc) rc = E->lanes.create();
    if (rc) { ens_unbind(E); E->lanes.destroy(); hipStreamDestroy(E->st); delete E; return rc; }
    *out = E;
    return 0;
}

int fdns_ens_step(fdns_ens* E, int nsteps, int use_graph)
{
    if (!E) return fail(FDNS_EINVAL, "null ensemble");
    if (E->kstep == 0)
        if (int rc = groups_retune(E->groups, E->tune_epoch, E->sg)) return rc;
    for (int k = 0; k < nsteps;) {
        if (int rc = ens_regroup(E)) return rc;
        const bool first = (E->kstep == 0);
        const bool rec = E->nKE > 0 && ((E->kstep + 1) % E->nKE) == 0;
        int n = 1;
        if (first || rec || !use_graph) {
            groups_step(E->groups, E->st, E->lanes, rec);
        } else {
            // the steady steps before the next record step: every group's
            // graph replayed n times on its lane between one fork and join
            n = nsteps - k;
            if (E->nKE > 0) n = std::min(n, E->nKE - 1 - E->kstep % E->nKE);
            if (E->sg.empty())
                if (int rc = E->sg.capture(E->groups, E->st, E->lanes)) return rc;
            if (int rc = E->sg.launch(E->st, E->lanes, n)) return rc;
        }
        E->kstep += n;
        for (auto& g : E->groups)
            for (auto* R : g.runs) R->kstep += n;
        k += n;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(E->st));
    return 0;
}

// Batched steps launched one kernel at a time with a HIP event pair around
// every launch, the groups one after another on the ensemble's stream.
// kernel_ms[c] / launches[c]: summed duration and count per kernel class
// FDNS_KC_*; wall_ms: host wall time of the call.
int fdns_ens_step_timed(fdns_ens* E, int nsteps, double* kernel_ms, int* launches, double* wall_ms)
{
    if (!E) return fail(FDNS_EINVAL, "null ensemble");
    if (E->kstep == 0)
        if (int rc = groups_retune(E->groups, E->tune_epoch, E->sg)) return rc;
    KTimer kt;
    HIPCHK(hipStreamSynchronize(E->st));
    const auto t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < nsteps; ++k) {
        if (int rc = ens_regroup(E)) return rc;
        const bool rec = E->nKE > 0 && ((E->kstep + 1) % E->nKE) == 0;
        groups_step(E->groups, E->st, E->lanes, rec, &kt);
        E->kstep++;
        for (auto& g : E->groups)
            for (auto* R : g.runs) R->kstep++;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(E->st));
    const auto t1 = std::chrono::steady_clock::now();
    if (wall_ms) *wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    double ms[KTimer::NCLS];
    int cnt[KTimer::NCLS];
    kt.collect(ms, cnt);
    for (int c = 0; c < KTimer::NCLS; ++c) {
        if (kernel_ms) kernel_ms[c] = ms[c];
        if (launches) launches[c] = cnt[c];
    }
    return 0;
}

int fdns_ens_destroy(fdns_ens* E)
{
    if (!E) return 0;
    hipStreamSynchronize(E->st);
    E->sg.release();
    ens_unbind(E);
    E->lanes.destroy();
    hipStreamDestroy(E->st);
    delete E;
    return 0;
}

int fdns_tuning_get(int Nr, int Nm, fdns_tuning* out)
{
    if (!out) return fail(FDNS_EINVAL, "null tuning");
    std::lock_guard<std::mutex> lk(g_mu);
    const auto key = std::make_pair(Nr, Nm);
    const auto it = g_shapes.find(key);
    if (it != g_shapes.end()) *out = tuning_of(it->second->sh);
    else {
        const auto tk = g_tune.find(key);
        *out = tk != g_tune.end() ? tk->second : tuning_default(Nr, Nm);
    }
    return 0;
}

int fdns_tuning_set(int Nr, int Nm, const fdns_tuning* t)
{
    if (!t || !tuning_valid(*t))
        return fail(FDNS_EINVAL, "tuning: nl_tpb 128|256, tile_waves[] 4|8, tile_stage[] and tile_pf[] 0|1, "
                                 "solve_chunk 0..32 in steps of 8, lane_split 0|1");
    std::lock_guard<std::mutex> lk(g_mu);
    g_tune[std::make_pair(Nr, Nm)] = *t;
    for (ShapeCtx* c : g_live)
        if (c->N == Nr && c->Nm == Nm) tuning_apply(c->sh, *t);
    g_tune_epoch++;
    return 0;
}

int fdns_ens_autotune(fdns_ens* E, int steps, fdns_tuning* chosen, double* report)
{
    if (!E || steps <= 0) return fail(FDNS_EINVAL, "autotune: ensemble and steps > 0");
    const int N = E->groups[0].shape->N, Nm = E->groups[0].shape->Nm;
    for (auto& g : E->groups)
        if (g.shape->N != N || g.shape->Nm != Nm) return fail(FDNS_EINVAL, "autotune: members of one (Nr_max, Nm_max)");
    fdns_tuning was;
    fdns_tuning_get(N, Nm, &was);
    // the options of each family; pass i takes option i mod count of each
    constexpr int NT = 8;   // tile options: (waves, stage, prefetch)
    const int nl[2] = {128, 256};
    int tw[NT], ts[NT], tp[NT];
    for (int o = 0; o < NT; ++o) { tw[o] = (o & 1) ? 8 : 4; ts[o] = (o >> 1) & 1; tp[o] = (o >> 2) & 1; }
    const int cap = solve_chunk_cap(N);
    int ch[3] = {0, 16, 8}, nch = 1;
    for (int c : {16, 8})
        if (c < cap) ch[nch++] = c;
    double t_nl[2] = {0, 0}, t_tile[6][NT] = {}, t_ch[3] = {0, 0, 0};
    int n_nl[2] = {0, 0}, n_ch[3] = {0, 0, 0}, n_tile[6][NT] = {};
    std::vector<fdns_run*> clones;
    fdns_ens* S = nullptr;
    int rc = 0;
    for (auto& g : E->groups) {
        for (auto* R : g.runs) {
            fdns_run* X = nullptr;
            if ((rc = fdns_create(&R->p, &X))) break;
            clones.push_back(X);
            if ((rc = fdns_seed_state(X, 7000 + clones.size(), 1e-3))) break;
        }
        if (rc) break;
    }
    if (!rc) rc = fdns_ens_create(clones.data(), (int)clones.size(), &S);
    if (!rc) rc = fdns_ens_step(S, 1, 0);   // the first step (prologue) untimed
    double t_split[2] = {-1.0, -1.0};
    for (int pass = 0; pass < NT && !rc; ++pass) {
        fdns_tuning t = was;
        t.nl_tpb = nl[pass % 2];
        for (int c = 0; c < 6; ++c) { t.tile_waves[c] = tw[pass]; t.tile_stage[c] = ts[pass]; t.tile_pf[c] = tp[pass]; }
        t.solve_chunk = ch[pass % nch];
        if ((rc = fdns_tuning_set(N, Nm, &t))) break;
        double ms[FDNS_NKCLASS];
        int cnt[FDNS_NKCLASS];
        if ((rc = fdns_ens_step_timed(S, 1, ms, cnt, nullptr))) break;   // warm-up with these choices
        if ((rc = fdns_ens_step_timed(S, steps, ms, cnt, nullptr))) break;
        t_nl[pass % 2] += ms[FDNS_KC_NONLIN] / steps;
        n_nl[pass % 2]++;
        for (int c = 0; c < 6; ++c)
            if (cnt[FDNS_KC_IMEX_PRE + c] > 0) {
                t_tile[c][pass] += ms[FDNS_KC_IMEX_PRE + c] / steps;
                n_tile[c][pass]++;
            }
        t_ch[pass % nch] += (ms[FDNS_KC_SOLVE_T] + ms[FDNS_KC_SOLVE_F]) / steps;
        n_ch[pass % nch]++;
    }
    // lane_split changes the concurrency of the groups, which the per-kernel
    // times above do not see: time graph-replayed steps with it off and on
    for (int sp = 0; sp < 2 && !rc; ++sp) {
        fdns_tuning t;
        if ((rc = fdns_tuning_get(N, Nm, &t))) break;
        t.lane_split = sp;
        if ((rc = fdns_tuning_set(N, Nm, &t))) break;
        if ((rc = fdns_ens_step(S, 2, 1))) break;   // regroup and capture
        const auto w0 = std::chrono::steady_clock::now();
        if ((rc = fdns_ens_step(S, steps, 1))) break;
        t_split[sp] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count() / steps;
    }
    fdns_tuning best = was;
    int b = 0;
    for (int i = 1; i < 2; ++i)
        if (t_nl[i] / n_nl[i] < t_nl[b] / n_nl[b]) b = i;
    best.nl_tpb = nl[b];
    for (int c = 0; c < 6; ++c) {
        if (n_tile[c][0] == 0) continue;   // not launched by this ensemble: keep
        b = 0;
        for (int i = 1; i < NT; ++i)
            if (t_tile[c][i] < t_tile[c][b]) b = i;
        best.tile_waves[c] = tw[b];
        best.tile_stage[c] = ts[b];
        best.tile_pf[c] = tp[b];
    }
    b = 0;
    for (int i = 1; i < nch; ++i)
        if (t_ch[i] / n_ch[i] < t_ch[b] / n_ch[b]) b = i;
    best.solve_chunk = ch[b];
    best.lane_split = (t_split[1] > 0.0 && t_split[1] < t_split[0]) ? 1 : 0;
    // the per-kernel times miss how the groups' kernels share the chip: time
    // graph-replayed steps with the new choice and with the one in force
    // before, and keep the faster
    double t_keep[2] = {-1.0, -1.0};   // [0] the choice before, [1] the new one
    for (int v = 0; v < 2 && !rc; ++v) {
        const fdns_tuning& t = v ? best : was;
        if ((rc = fdns_tuning_set(N, Nm, &t))) break;
        if ((rc = fdns_ens_step(S, 2, 1))) break;   // regroup and capture
        const int n = 4 * steps;
        const auto w0 = std::chrono::steady_clock::now();
        if ((rc = fdns_ens_step(S, n, 1))) break;
        t_keep[v] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count() / n;
    }
    if (S) fdns_ens_destroy(S);
    for (auto* X : clones) fdns_destroy(X);
    if (rc) {
        fdns_tuning_set(N, Nm, &was);
        return rc;
    }
    if (t_keep[0] < t_keep[1]) best = was;
    if ((rc = fdns_tuning_set(N, Nm, &best))) return rc;
    if (chosen) *chosen = best;
    if (report) {
        for (int i = 0; i < FDNS_TUNE_REPORT; ++i) report[i] = -1.0;
        for (int i = 0; i < 2; ++i) report[i] = t_nl[i] / n_nl[i];
        for (int c = 0; c < 6; ++c)
            for (int i = 0; i < NT; ++i)
                if (n_tile[c][i] > 0) report[2 + NT * c + i] = t_tile[c][i];
        for (int i = 0; i < nch; ++i) report[2 + 6 * NT + (ch[i] == 0 ? 0 : (ch[i] == 16 ? 1 : 2))] = t_ch[i] / n_ch[i];
        report[2 + 6 * NT + 3] = t_split[0];
        report[2 + 6 * NT + 4] = t_split[1];
        report[2 + 6 * NT + 5] = t_keep[0];
        report[2 + 6 * NT + 6] = t_keep[1];
    }
    return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ serving
// A pool of simulation jobs.  Jobs wait in FIFO order until their footprint
// (state plus any factor set not yet resident) fits the HBM budget; resident
// jobs are regrouped by shape and scheme structure and advanced together, one
// batched step per fdns_pool_step; a finished job's state is copied to host
// memory and its device memory freed, which lets waiting jobs in.
namespace {

struct Job {
    int id = 0;
    fdns_params p;
    uint64_t seed = 0;
    double amp = 0.0;
    int nsteps = 0;
    int status = 0;          // 0 waiting, 1 running, 2 done
    int done = 0;
    fdns_run* run = nullptr;
    std::vector<double> tFR, omgFR, urFR, upFR;
    double diag[6] = {0};
};


}  // namespace

struct fdns_pool {
    size_t budget = 0;
    int next_id = 1;
    std::map<int, Job> jobs;
    std::vector<int> waiting;   // FIFO
    std::vector<int> active;
    std::vector<Group> groups;  // resident runs past their first step, by structure
    Lanes lanes;
    hipStream_t st = nullptr;
    StepGraphs sg;                   // one batched step of every group
    bool recapture = true;           // the groups changed since the capture
    int tune_epoch = 0;              // launch choices of the solve plans and the graph
};

namespace {

// Give group g the member list `runs`: its old members go home to their own
// descriptors, the new list gets one contiguous descriptor array (copied from
// each run's current descriptor, so device-side counters carry over) and a
// solve plan.  Other groups are not touched.
int group_set(fdns_pool* P, Group& g, const std::vector<fdns_run*>& runs)
{
    for (size_t i = 0; i < g.runs.size(); ++i) {
        fdns_run* R = g.runs[i];
        HIPCHK(hipMemcpyAsync(R->own_dev, g.dev + i, sizeof(RunDesc), hipMemcpyDeviceToDevice, P->st));
        R->dev = R->own_dev;
        R->in_ens = false;
    }
    HIPCHK(hipStreamSynchronize(P->st));
    if (g.dev) hipFree(g.dev);
    g.dev = nullptr;
    g.solve.release();
    g.runs = runs;
    if (runs.empty()) return 0;
    g.keep = false;
    for (auto* R : runs) g.keep = g.keep || R->h.keep;
    const int n = (int)runs.size();
    if (hipMalloc((void**)&g.dev, sizeof(RunDesc) * n) != hipSuccess) return fail(FDNS_ENOMEM, "pool descriptors");
    for (int i = 0; i < n; ++i) {
        fdns_run* R = runs[i];
        HIPCHK(hipMemcpyAsync(g.dev + i, R->dev, sizeof(RunDesc), hipMemcpyDeviceToDevice, P->st));
        R->dev = g.dev + i;
        R->in_ens = true;
    }
    int rc = plan_build(g.runs, g.solve);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(P->st));
    return 0;
}

// A run past its first step joins the group of its structure (a new group if
// none matches).
int pool_join(fdns_pool* P, fdns_run* R)
{
    for (auto& g : P->groups)
        if (same_structure(g.runs[0], R)) {
            std::vector<fdns_run*> r = g.runs;
            r.push_back(R);
            P->recapture = true;
            return group_set(P, g, r);
        }
    Group g;
    g.shape = R->shape; g.sc = R->sc; g.restart = R->p.l_restart != 0;
    P->groups.push_back(g);
    P->recapture = true;
    return group_set(P, P->groups.back(), {R});
}

// Runs leave their groups; emptied groups are dropped.
int pool_leave(fdns_pool* P, const std::set<fdns_run*>& gone)
{
    int rc = 0;
    for (auto& g : P->groups) {
        std::vector<fdns_run*> keep;
        for (auto* R : g.runs)
            if (!gone.count(R)) keep.push_back(R);
        if (keep.size() != g.runs.size()) {
            P->recapture = true;
            if ((rc = group_set(P, g, keep))) return rc;
        }
    }
    P->groups.erase(std::remove_if(P->groups.begin(), P->groups.end(), [](const Group& g) { return g.runs.empty(); }),
                    P->groups.end());
    return 0;
}

void pool_unbind(fdns_pool* P)
{
    for (auto& g : P->groups) group_set(P, g, {});
    P->groups.clear();
    P->sg.release();
    P->recapture = true;
}

// Device bytes held by the resident jobs: their states and each operator set
// they use, counted once however many jobs share it.
size_t pool_used(fdns_pool* P)
{
    std::set<const OpSet*> sets;
    size_t b = 0;
    for (int id : P->active) {
        const Job& J = P->jobs[id];
        b += J.run->bytes;
        const fdns_run* R = J.run;
        for (const OpSet* o : {R->opT, R->opF, R->opL}) sets.insert(o);
        for (int k = 0; k < MAXS; ++k) {
            sets.insert(R->stT[k]);
            sets.insert(R->stF[k]);
        }
    }
    sets.erase(nullptr);
    for (const OpSet* S : sets) b += S->bytes;
    return b;
}

// Bytes admitting J adds: its state and the operator sets no resident job uses.
size_t job_need(fdns_pool* P, const Job& J)
{
    // the job's whole footprint (state, every operator set it binds), less
    // the main temperature and vorticity sets a resident job already holds
    Scheme sc;
    scheme_lookup(J.p.time_scheme_type, J.p.time_scheme_imp, J.p.time_scheme_exp, sc);
    size_t b = fdns_job_footprint(&J.p);
    if (sc.kind == SK_RK) return b;
    OpKey kT, kF;
    run_keys(&J.p, sc, kT, kF);
    bool hasT = false, hasF = false;
    for (int id : P->active) {
        const fdns_run* R = P->jobs[id].run;
        hasT = hasT || (R->opT && R->opT->key == kT);
        hasF = hasF || (R->opF && R->opF->key == kF);
    }
    if (hasT) b -= opset_bytes(kT.kind, J.p.Nr, J.p.Nm);
    if (hasF) b -= opset_bytes(kF.kind, J.p.Nr, J.p.Nm);
    return b;
}

// One call of the serving loop: admit the waiting jobs that fit (FIFO), give
// each its first step on its own, advance the resident steady-state jobs by k
// batched steps -- replays of one captured graph of every group, captured
// again only when the groups changed -- with k the smaller of max_steps and
// the fewest steps any resident job has left, synchronise once, and retire
// the jobs that are done (their state to host memory, their device memory
// freed).  New jobs join their groups for the next call.
int pool_run(fdns_pool* P, int max_steps, int* taken)
{
    int rc;
    std::vector<int> fresh;
    while (!P->waiting.empty()) {
        Job& J = P->jobs[P->waiting.front()];
        if (pool_used(P) + job_need(P, J) > P->budget) break;
        fdns_run* R = nullptr;
        if ((rc = fdns_create(&J.p, &R))) return rc;
        if ((rc = fdns_seed_state(R, J.seed, J.amp))) { fdns_destroy(R); return rc; }
        J.run = R;
        J.status = 1;
        P->active.push_back(J.id);
        P->waiting.erase(P->waiting.begin());
        fresh.push_back(J.id);
    }
    // the steady-state jobs' batch length
    int k = max_steps;
    for (int id : P->active) {
        const Job& J = P->jobs[id];
        if (std::find(fresh.begin(), fresh.end(), id) == fresh.end()) k = std::min(k, J.nsteps - J.done);
    }
    const bool batch = !P->groups.empty() && k > 0;
    if (P->tune_epoch != g_tune_epoch) {
        if (int rc = groups_retune(P->groups, P->tune_epoch, P->sg)) return rc;
        P->recapture = true;
    }
    if (batch && P->recapture) {
        if (int rc = P->sg.capture(P->groups, P->st, P->lanes)) return rc;
        P->recapture = false;
    }
    // first steps of the new jobs, each with its own solve plan
    for (int id : fresh) {
        fdns_run* R = P->jobs[id].run;
        const Shape& sh = R->shape->sh;
        const bool restart = R->p.l_restart != 0;
        prologue_launches(R->dev, 1, sh, R->sc, restart, P->st);
        step_launches(R->dev, 1, sh, R->sc, R->solo, true, restart, false, P->st, nullptr, R->h.keep != 0, &R->side);
    }
    if (batch)
        if (int rc = P->sg.launch(P->st, P->lanes, k)) return rc;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(P->st));
    for (int id : P->active) {
        Job& J = P->jobs[id];
        const bool is_new = std::find(fresh.begin(), fresh.end(), id) != fresh.end();
        const int n = is_new ? 1 : (batch ? k : 0);
        J.done += n;
        J.run->kstep += n;
    }
    if (taken) *taken = batch ? k : (fresh.empty() ? 0 : 1);
    // retire, then let the new jobs that go on join their groups
    std::set<fdns_run*> gone;
    std::vector<int> keep;
    for (int id : P->active) {
        Job& J = P->jobs[id];
        if (J.done >= J.nsteps) gone.insert(J.run);
        else keep.push_back(id);
    }
    if (!gone.empty() && (rc = pool_leave(P, gone))) return rc;
    for (int id : P->active) {
        Job& J = P->jobs[id];
        if (!gone.count(J.run)) continue;
        const size_t S = (size_t)(J.p.Nm + 1) * J.p.Nr;
        J.tFR.resize(2 * S); J.omgFR.resize(2 * S); J.urFR.resize(2 * S); J.upFR.resize(2 * S);
        if ((rc = fdns_get_state(J.run, J.tFR.data(), J.omgFR.data(), J.urFR.data(), J.upFR.data(), nullptr))) return rc;
        if ((rc = fdns_diag(J.run, J.diag))) return rc;
        fdns_destroy(J.run);
        J.run = nullptr;
        J.status = 2;
    }
    P->active = keep;
    for (int id : fresh) {
        Job& J = P->jobs[id];
        if (J.status == 1 && (rc = pool_join(P, J.run))) return rc;
    }
    return 0;
}

}  // namespace

extern "C" {

int fdns_pool_create(size_t hbm_budget_bytes, fdns_pool** out)
{
    if (!out) return fail(FDNS_EINVAL, "null out");
    int rc;
    if ((rc = ensure_kernels())) return rc;
    auto* P = new fdns_pool();
    P->budget = hbm_budget_bytes;
    P->tune_epoch = g_tune_epoch;
    if (hipStreamCreateWithFlags(&P->st, hipStreamNonBlocking) != hipSuccess) { delete P; return fail(FDNS_EHIP, "stream"); }
    if ((rc = P->lanes.create())) { P->lanes.destroy(); hipStreamDestroy(P->st); delete P; return rc; }
    *out = P;
    return 0;
}

int fdns_pool_submit(fdns_pool* P, const fdns_params* p, uint64_t seed, double amp, int nsteps, int* job_id)
{
    if (!P || !p || !job_id || nsteps <= 0) return fail(FDNS_EINVAL, "bad submit arguments");
    Scheme sc;
    int rc = validate(p, sc);
    if (rc) return rc;
    if (p->l_vartimestep) return fail(FDNS_EINVAL, "variable-dt runs step alone, not in a pool");
    Job J;
    J.id = P->next_id++;
    J.p = *p;
    J.seed = seed;
    J.amp = amp;
    J.nsteps = nsteps;
    if (fdns_job_footprint(p) > P->budget) return fail(FDNS_EBUSY, "job larger than the pool budget");
    P->jobs[J.id] = J;
    P->waiting.push_back(J.id);
    *job_id = J.id;
    return 0;
}

int fdns_pool_step(fdns_pool* P, int* n_active, int* n_waiting)
{
    if (!P) return fail(FDNS_EINVAL, "null pool");
    int rc = pool_run(P, 1, nullptr);
    if (rc) return rc;
    if (n_active) *n_active = (int)P->active.size();
    if (n_waiting) *n_waiting = (int)P->waiting.size();
    return 0;
}

int fdns_pool_run(fdns_pool* P, int max_steps, int* steps_taken, int* n_active, int* n_waiting)
{
    if (!P || max_steps < 1) return fail(FDNS_EINVAL, "null pool or max_steps < 1");
    int rc = pool_run(P, max_steps, steps_taken);
    if (rc) return rc;
    if (n_active) *n_active = (int)P->active.size();
    if (n_waiting) *n_waiting = (int)P->waiting.size();
    return 0;
}

int fdns_pool_poll(fdns_pool* P, int job_id, int* status, int* steps_done)
{
    if (!P) return fail(FDNS_EINVAL, "null pool");
    auto it = P->jobs.find(job_id);
    if (it == P->jobs.end()) return fail(FDNS_ENOTFOUND, "unknown job");
    if (status) *status = it->second.status;
    if (steps_done) *steps_done = it->second.done;
    return 0;
}

int fdns_pool_result(fdns_pool* P, int job_id, double* tFR, double* omgFR, double* urFR, double* upFR, double* diag6)
{
    if (!P) return fail(FDNS_EINVAL, "null pool");
    auto it = P->jobs.find(job_id);
    if (it == P->jobs.end()) return fail(FDNS_ENOTFOUND, "unknown job");
    Job& J = it->second;
    if (J.status != 2) return fail(FDNS_ESTATE, "job not finished");
    const size_t n = J.tFR.size();
    if (tFR) std::memcpy(tFR, J.tFR.data(), sizeof(double) * n);
    if (omgFR) std::memcpy(omgFR, J.omgFR.data(), sizeof(double) * n);
    if (urFR) std::memcpy(urFR, J.urFR.data(), sizeof(double) * n);
    if (upFR) std::memcpy(upFR, J.upFR.data(), sizeof(double) * n);
    if (diag6) std::memcpy(diag6, J.diag, sizeof(J.diag));
    return 0;
}

int fdns_pool_release(fdns_pool* P, int job_id)
{
    if (!P) return fail(FDNS_EINVAL, "null pool");
    auto it = P->jobs.find(job_id);
    if (it == P->jobs.end()) return fail(FDNS_ENOTFOUND, "unknown job");
    if (it->second.status == 1) return fail(FDNS_ESTATE, "job is running");
    if (it->second.status == 0) P->waiting.erase(std::find(P->waiting.begin(), P->waiting.end(), job_id));
    P->jobs.erase(it);
    return 0;
}

int fdns_pool_destroy(fdns_pool* P)
{
    if (!P) return 0;
    pool_unbind(P);
    for (auto& kv : P->jobs)
        if (kv.second.run) fdns_destroy(kv.second.run);
    P->lanes.destroy();
    hipStreamDestroy(P->st);
    delete P;
    return 0;
}

}  // extern "C"
