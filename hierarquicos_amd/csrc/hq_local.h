// Generated by gen_local.py from the formulas in its docstring; do not edit.
#pragma once

namespace hq {

constexpr int NLOC = 55;  // local coefficients, 1 <= |n| <= 5
// LOC_n[i] = (nx, ny, nz): 100 010 001 200 110 101 020 011 002 300 210 201 120 111 102 030 021 012 003 400 310 301 220 211 202 130 121 112 103 040 031 022 013 004 500 410 401 320 311 302 230 221 212 203 140 131 122 113 104 050 041 032 023 014 005

// Local coefficients of one source cell (this lane's) about the bunch centre.
// R = bunch centre - cell centre of mass, h = bunch radius (> 0), r = the
// cell's FP32 record; adds h^|n| L_n to L[i] for the multi-index n = LOC_n[i]
// (the addition folded into each coefficient's last multiply-add).
template <int MP>
__device__ __forceinline__ void local_m2l(float Rx, float Ry, float Rz, float eps2, float h,
                                          const float* __restrict__ r, float* __restrict__ L) {
#pragma clang fp contract(off)
    const float rho = fmaf(Rx, Rx, fmaf(Ry, Ry, fmaf(Rz, Rz, eps2)));
    const float ri = __frsqrt_rn(rho);
    const float X = Rx * ri;
    const float Y = Ry * ri;
    const float Z = Rz * ri;
    const float sg = h * ri;
    const float m200 = X * X;
    const float m110 = Y * X;
    const float m101 = Z * X;
    const float m020 = Y * Y;
    const float m011 = Z * Y;
    const float m002 = Z * Z;
    const float m300 = m200 * X;
    const float m210 = m110 * X;
    const float m201 = m101 * X;
    const float m120 = m020 * X;
    const float m111 = m011 * X;
    const float m102 = m002 * X;
    const float m030 = m020 * Y;
    const float m021 = m011 * Y;
    const float m012 = m002 * Y;
    const float m003 = m002 * Z;
    const float m400 = m300 * X;
    const float m310 = m210 * X;
    const float m301 = m201 * X;
    const float m220 = m120 * X;
    const float m211 = m111 * X;
    const float m202 = m102 * X;
    const float m130 = m030 * X;
    const float m121 = m021 * X;
    const float m112 = m012 * X;
    const float m103 = m003 * X;
    const float m040 = m030 * Y;
    const float m031 = m021 * Y;
    const float m022 = m012 * Y;
    const float m013 = m003 * Y;
    const float m004 = m003 * Z;
    const float m500 = m400 * X;
    const float m410 = m310 * X;
    const float m401 = m301 * X;
    const float m320 = m220 * X;
    const float m311 = m211 * X;
    const float m302 = m202 * X;
    const float m230 = m130 * X;
    const float m221 = m121 * X;
    const float m212 = m112 * X;
    const float m203 = m103 * X;
    const float m140 = m040 * X;
    const float m131 = m031 * X;
    const float m122 = m022 * X;
    const float m113 = m013 * X;
    const float m104 = m004 * X;
    const float m050 = m040 * Y;
    const float m041 = m031 * Y;
    const float m032 = m022 * Y;
    const float m023 = m013 * Y;
    const float m014 = m004 * Y;
    const float m005 = m004 * Z;
    const float D100 = -X;
    const float D010 = -Y;
    const float D001 = -Z;
    const float D200 = fmaf(3.0f, m200, -1.0f);
    const float D110 = 3.0f * m110;
    const float D101 = 3.0f * m101;
    const float D020 = fmaf(3.0f, m020, -1.0f);
    const float D011 = 3.0f * m011;
    const float D002 = fmaf(3.0f, m002, -1.0f);
    const float D300 = fmaf(9.0f, X, -15.0f * m300);
    const float D210 = fmaf(3.0f, Y, -15.0f * m210);
    const float D201 = fmaf(3.0f, Z, -15.0f * m201);
    const float D120 = fmaf(3.0f, X, -15.0f * m120);
    const float D111 = -15.0f * m111;
    const float D102 = fmaf(3.0f, X, -15.0f * m102);
    const float D030 = fmaf(9.0f, Y, -15.0f * m030);
    const float D021 = fmaf(3.0f, Z, -15.0f * m021);
    const float D012 = fmaf(3.0f, Y, -15.0f * m012);
    const float D003 = fmaf(9.0f, Z, -15.0f * m003);
    const float D400 = fmaf(-90.0f, m200, fmaf(105.0f, m400, 9.0f));
    const float D310 = fmaf(-45.0f, m110, 105.0f * m310);
    const float D301 = fmaf(-45.0f, m101, 105.0f * m301);
    const float D220 = fmaf(-15.0f, m200, fmaf(-15.0f, m020, fmaf(105.0f, m220, 3.0f)));
    const float D211 = fmaf(-15.0f, m011, 105.0f * m211);
    const float D202 = fmaf(-15.0f, m200, fmaf(-15.0f, m002, fmaf(105.0f, m202, 3.0f)));
    const float D130 = fmaf(-45.0f, m110, 105.0f * m130);
    const float D121 = fmaf(-15.0f, m101, 105.0f * m121);
    const float D112 = fmaf(-15.0f, m110, 105.0f * m112);
    const float D103 = fmaf(-45.0f, m101, 105.0f * m103);
    const float D040 = fmaf(-90.0f, m020, fmaf(105.0f, m040, 9.0f));
    const float D031 = fmaf(-45.0f, m011, 105.0f * m031);
    const float D022 = fmaf(-15.0f, m020, fmaf(-15.0f, m002, fmaf(105.0f, m022, 3.0f)));
    const float D013 = fmaf(-45.0f, m011, 105.0f * m013);
    const float D004 = fmaf(-90.0f, m002, fmaf(105.0f, m004, 9.0f));
    const float D500 = fmaf(-225.0f, X, fmaf(1050.0f, m300, -945.0f * m500));
    const float D410 = fmaf(-45.0f, Y, fmaf(630.0f, m210, -945.0f * m410));
    const float D401 = fmaf(-45.0f, Z, fmaf(630.0f, m201, -945.0f * m401));
    const float D320 = fmaf(-45.0f, X, fmaf(105.0f, m300, fmaf(315.0f, m120, -945.0f * m320)));
    const float D311 = fmaf(315.0f, m111, -945.0f * m311);
    const float D302 = fmaf(-45.0f, X, fmaf(105.0f, m300, fmaf(315.0f, m102, -945.0f * m302)));
    const float D230 = fmaf(-45.0f, Y, fmaf(315.0f, m210, fmaf(105.0f, m030, -945.0f * m230)));
    const float D221 = fmaf(-15.0f, Z, fmaf(105.0f, m201, fmaf(105.0f, m021, -945.0f * m221)));
    const float D212 = fmaf(-15.0f, Y, fmaf(105.0f, m210, fmaf(105.0f, m012, -945.0f * m212)));
    const float D203 = fmaf(-45.0f, Z, fmaf(315.0f, m201, fmaf(105.0f, m003, -945.0f * m203)));
    const float D140 = fmaf(-45.0f, X, fmaf(630.0f, m120, -945.0f * m140));
    const float D131 = fmaf(315.0f, m111, -945.0f * m131);
    const float D122 = fmaf(-15.0f, X, fmaf(105.0f, m120, fmaf(105.0f, m102, -945.0f * m122)));
    const float D113 = fmaf(315.0f, m111, -945.0f * m113);
    const float D104 = fmaf(-45.0f, X, fmaf(630.0f, m102, -945.0f * m104));
    const float D050 = fmaf(-225.0f, Y, fmaf(1050.0f, m030, -945.0f * m050));
    const float D041 = fmaf(-45.0f, Z, fmaf(630.0f, m021, -945.0f * m041));
    const float D032 = fmaf(-45.0f, Y, fmaf(105.0f, m030, fmaf(315.0f, m012, -945.0f * m032)));
    const float D023 = fmaf(-45.0f, Z, fmaf(315.0f, m021, fmaf(105.0f, m003, -945.0f * m023)));
    const float D014 = fmaf(-45.0f, Y, fmaf(630.0f, m012, -945.0f * m014));
    const float D005 = fmaf(-225.0f, Z, fmaf(1050.0f, m003, -945.0f * m005));
    const float ri2 = ri * ri;
    const float ri3 = ri2 * ri;
    const float P000 = -r[0];
    const float cq0 = MP >= 4 ? ri2 * 0.16666666666666666f : 0.0f;
    const float cq1 = MP >= 4 ? ri2 * 0.3333333333333333f : 0.0f;
    const float co0 = MP == 8 ? ri3 * 0.022222222222222223f : 0.0f;
    const float co1 = MP == 8 ? ri3 * 0.06666666666666667f : 0.0f;
    const float co2 = MP == 8 ? ri3 * 0.13333333333333333f : 0.0f;
    const float P200 = MP >= 4 ? r[1] * cq0 : 0.0f;
    const float P020 = MP >= 4 ? r[2] * cq0 : 0.0f;
    const float P002 = MP >= 4 ? r[3] * cq0 : 0.0f;
    const float P110 = MP >= 4 ? r[4] * cq1 : 0.0f;
    const float P101 = MP >= 4 ? r[5] * cq1 : 0.0f;
    const float P011 = MP >= 4 ? r[6] * cq1 : 0.0f;
    const float P300 = MP == 8 ? r[11] * co0 : 0.0f;
    const float P030 = MP == 8 ? r[15] * co0 : 0.0f;
    const float P003 = MP == 8 ? r[19] * co0 : 0.0f;
    const float P120 = MP == 8 ? r[12] * co1 : 0.0f;
    const float P102 = MP == 8 ? r[13] * co1 : 0.0f;
    const float P210 = MP == 8 ? r[14] * co1 : 0.0f;
    const float P012 = MP == 8 ? r[16] * co1 : 0.0f;
    const float P201 = MP == 8 ? r[17] * co1 : 0.0f;
    const float P021 = MP == 8 ? r[18] * co1 : 0.0f;
    const float P111 = MP == 8 ? r[20] * co2 : 0.0f;
    const float s1 = ri * sg;
    const float s2 = s1 * sg;
    const float s3 = s2 * sg;
    const float s4 = s3 * sg;
    const float s5 = s4 * sg;
    const float q0 = P000 * ri;
    const float t1 = q0 * sg;
    const float t2 = t1 * sg;
    const float t3 = t2 * sg;
    const float t4 = t3 * sg;
    const float t5 = t4 * sg;
    L[0] = fmaf(s1, fmaf(P111, D211, fmaf(P003, D103, fmaf(P021, D121, fmaf(P201, D301, fmaf(P012, D112, fmaf(P030, D130, fmaf(P210, D310, fmaf(P102, D202, fmaf(P120, D220, fmaf(P300, D400, fmaf(P011, D111, fmaf(P101, D201, fmaf(P110, D210, fmaf(P002, D102, fmaf(P020, D120, P200 * D300))))))))))))))), fmaf(t1, D100, L[0]));
    L[1] = fmaf(s1, fmaf(P111, D121, fmaf(P003, D013, fmaf(P021, D031, fmaf(P201, D211, fmaf(P012, D022, fmaf(P030, D040, fmaf(P210, D220, fmaf(P102, D112, fmaf(P120, D130, fmaf(P300, D310, fmaf(P011, D021, fmaf(P101, D111, fmaf(P110, D120, fmaf(P002, D012, fmaf(P020, D030, P200 * D210))))))))))))))), fmaf(t1, D010, L[1]));
    L[2] = fmaf(s1, fmaf(P111, D112, fmaf(P003, D004, fmaf(P021, D022, fmaf(P201, D202, fmaf(P012, D013, fmaf(P030, D031, fmaf(P210, D211, fmaf(P102, D103, fmaf(P120, D121, fmaf(P300, D301, fmaf(P011, D012, fmaf(P101, D102, fmaf(P110, D111, fmaf(P002, D003, fmaf(P020, D021, P200 * D201))))))))))))))), fmaf(t1, D001, L[2]));
    L[3] = fmaf(s2, fmaf(P111, D311, fmaf(P003, D203, fmaf(P021, D221, fmaf(P201, D401, fmaf(P012, D212, fmaf(P030, D230, fmaf(P210, D410, fmaf(P102, D302, fmaf(P120, D320, fmaf(P300, D500, fmaf(P011, D211, fmaf(P101, D301, fmaf(P110, D310, fmaf(P002, D202, fmaf(P020, D220, P200 * D400))))))))))))))), fmaf(t2, D200, L[3]));
    L[4] = fmaf(s2, fmaf(P111, D221, fmaf(P003, D113, fmaf(P021, D131, fmaf(P201, D311, fmaf(P012, D122, fmaf(P030, D140, fmaf(P210, D320, fmaf(P102, D212, fmaf(P120, D230, fmaf(P300, D410, fmaf(P011, D121, fmaf(P101, D211, fmaf(P110, D220, fmaf(P002, D112, fmaf(P020, D130, P200 * D310))))))))))))))), fmaf(t2, D110, L[4]));
    L[5] = fmaf(s2, fmaf(P111, D212, fmaf(P003, D104, fmaf(P021, D122, fmaf(P201, D302, fmaf(P012, D113, fmaf(P030, D131, fmaf(P210, D311, fmaf(P102, D203, fmaf(P120, D221, fmaf(P300, D401, fmaf(P011, D112, fmaf(P101, D202, fmaf(P110, D211, fmaf(P002, D103, fmaf(P020, D121, P200 * D301))))))))))))))), fmaf(t2, D101, L[5]));
    L[6] = fmaf(s2, fmaf(P111, D131, fmaf(P003, D023, fmaf(P021, D041, fmaf(P201, D221, fmaf(P012, D032, fmaf(P030, D050, fmaf(P210, D230, fmaf(P102, D122, fmaf(P120, D140, fmaf(P300, D320, fmaf(P011, D031, fmaf(P101, D121, fmaf(P110, D130, fmaf(P002, D022, fmaf(P020, D040, P200 * D220))))))))))))))), fmaf(t2, D020, L[6]));
    L[7] = fmaf(s2, fmaf(P111, D122, fmaf(P003, D014, fmaf(P021, D032, fmaf(P201, D212, fmaf(P012, D023, fmaf(P030, D041, fmaf(P210, D221, fmaf(P102, D113, fmaf(P120, D131, fmaf(P300, D311, fmaf(P011, D022, fmaf(P101, D112, fmaf(P110, D121, fmaf(P002, D013, fmaf(P020, D031, P200 * D211))))))))))))))), fmaf(t2, D011, L[7]));
    L[8] = fmaf(s2, fmaf(P111, D113, fmaf(P003, D005, fmaf(P021, D023, fmaf(P201, D203, fmaf(P012, D014, fmaf(P030, D032, fmaf(P210, D212, fmaf(P102, D104, fmaf(P120, D122, fmaf(P300, D302, fmaf(P011, D013, fmaf(P101, D103, fmaf(P110, D112, fmaf(P002, D004, fmaf(P020, D022, P200 * D202))))))))))))))), fmaf(t2, D002, L[8]));
    L[9] = fmaf(s3, fmaf(P011, D311, fmaf(P101, D401, fmaf(P110, D410, fmaf(P002, D302, fmaf(P020, D320, P200 * D500))))), fmaf(t3, D300, L[9]));
    L[10] = fmaf(s3, fmaf(P011, D221, fmaf(P101, D311, fmaf(P110, D320, fmaf(P002, D212, fmaf(P020, D230, P200 * D410))))), fmaf(t3, D210, L[10]));
    L[11] = fmaf(s3, fmaf(P011, D212, fmaf(P101, D302, fmaf(P110, D311, fmaf(P002, D203, fmaf(P020, D221, P200 * D401))))), fmaf(t3, D201, L[11]));
    L[12] = fmaf(s3, fmaf(P011, D131, fmaf(P101, D221, fmaf(P110, D230, fmaf(P002, D122, fmaf(P020, D140, P200 * D320))))), fmaf(t3, D120, L[12]));
    L[13] = fmaf(s3, fmaf(P011, D122, fmaf(P101, D212, fmaf(P110, D221, fmaf(P002, D113, fmaf(P020, D131, P200 * D311))))), fmaf(t3, D111, L[13]));
    L[14] = fmaf(s3, fmaf(P011, D113, fmaf(P101, D203, fmaf(P110, D212, fmaf(P002, D104, fmaf(P020, D122, P200 * D302))))), fmaf(t3, D102, L[14]));
    L[15] = fmaf(s3, fmaf(P011, D041, fmaf(P101, D131, fmaf(P110, D140, fmaf(P002, D032, fmaf(P020, D050, P200 * D230))))), fmaf(t3, D030, L[15]));
    L[16] = fmaf(s3, fmaf(P011, D032, fmaf(P101, D122, fmaf(P110, D131, fmaf(P002, D023, fmaf(P020, D041, P200 * D221))))), fmaf(t3, D021, L[16]));
    L[17] = fmaf(s3, fmaf(P011, D023, fmaf(P101, D113, fmaf(P110, D122, fmaf(P002, D014, fmaf(P020, D032, P200 * D212))))), fmaf(t3, D012, L[17]));
    L[18] = fmaf(s3, fmaf(P011, D014, fmaf(P101, D104, fmaf(P110, D113, fmaf(P002, D005, fmaf(P020, D023, P200 * D203))))), fmaf(t3, D003, L[18]));
    L[19] = fmaf(t4, D400, L[19]);
    L[20] = fmaf(t4, D310, L[20]);
    L[21] = fmaf(t4, D301, L[21]);
    L[22] = fmaf(t4, D220, L[22]);
    L[23] = fmaf(t4, D211, L[23]);
    L[24] = fmaf(t4, D202, L[24]);
    L[25] = fmaf(t4, D130, L[25]);
    L[26] = fmaf(t4, D121, L[26]);
    L[27] = fmaf(t4, D112, L[27]);
    L[28] = fmaf(t4, D103, L[28]);
    L[29] = fmaf(t4, D040, L[29]);
    L[30] = fmaf(t4, D031, L[30]);
    L[31] = fmaf(t4, D022, L[31]);
    L[32] = fmaf(t4, D013, L[32]);
    L[33] = fmaf(t4, D004, L[33]);
    L[34] = fmaf(t5, D500, L[34]);
    L[35] = fmaf(t5, D410, L[35]);
    L[36] = fmaf(t5, D401, L[36]);
    L[37] = fmaf(t5, D320, L[37]);
    L[38] = fmaf(t5, D311, L[38]);
    L[39] = fmaf(t5, D302, L[39]);
    L[40] = fmaf(t5, D230, L[40]);
    L[41] = fmaf(t5, D221, L[41]);
    L[42] = fmaf(t5, D212, L[42]);
    L[43] = fmaf(t5, D203, L[43]);
    L[44] = fmaf(t5, D140, L[44]);
    L[45] = fmaf(t5, D131, L[45]);
    L[46] = fmaf(t5, D122, L[46]);
    L[47] = fmaf(t5, D113, L[47]);
    L[48] = fmaf(t5, D104, L[48]);
    L[49] = fmaf(t5, D050, L[49]);
    L[50] = fmaf(t5, D041, L[50]);
    L[51] = fmaf(t5, D032, L[51]);
    L[52] = fmaf(t5, D023, L[52]);
    L[53] = fmaf(t5, D014, L[53]);
    L[54] = fmaf(t5, D005, L[54]);
}

// Acceleration sum_n L_(n+e_i) u^n / n! at u = (target - bunch centre) / h from the
// scaled coefficients L of local_m2l; the caller multiplies by -1/h.
__device__ __forceinline__ void local_l2p(const float* __restrict__ L, float ux, float uy,
                                          float uz, float acc[3]) {
#pragma clang fp contract(off)
    const float u200 = (ux * ux) * 0.5f;
    const float u110 = uy * ux;
    const float u101 = uz * ux;
    const float u020 = (uy * uy) * 0.5f;
    const float u011 = uz * uy;
    const float u002 = (uz * uz) * 0.5f;
    const float u300 = (u200 * ux) * 0.3333333333333333f;
    const float u210 = (u110 * ux) * 0.5f;
    const float u201 = (u101 * ux) * 0.5f;
    const float u120 = u020 * ux;
    const float u111 = u011 * ux;
    const float u102 = u002 * ux;
    const float u030 = (u020 * uy) * 0.3333333333333333f;
    const float u021 = (u011 * uy) * 0.5f;
    const float u012 = u002 * uy;
    const float u003 = (u002 * uz) * 0.3333333333333333f;
    const float u400 = (u300 * ux) * 0.25f;
    const float u310 = (u210 * ux) * 0.3333333333333333f;
    const float u301 = (u201 * ux) * 0.3333333333333333f;
    const float u220 = (u120 * ux) * 0.5f;
    const float u211 = (u111 * ux) * 0.5f;
    const float u202 = (u102 * ux) * 0.5f;
    const float u130 = u030 * ux;
    const float u121 = u021 * ux;
    const float u112 = u012 * ux;
    const float u103 = u003 * ux;
    const float u040 = (u030 * uy) * 0.25f;
    const float u031 = (u021 * uy) * 0.3333333333333333f;
    const float u022 = (u012 * uy) * 0.5f;
    const float u013 = u003 * uy;
    const float u004 = (u003 * uz) * 0.25f;
    acc[0] = fmaf(L[5], uz, fmaf(L[4], uy, fmaf(L[3], ux, fmaf(L[14], u002, fmaf(L[13], u011, fmaf(L[12], u020, fmaf(L[11], u101, fmaf(L[10], u110, fmaf(L[9], u200, fmaf(L[28], u003, fmaf(L[27], u012, fmaf(L[26], u021, fmaf(L[25], u030, fmaf(L[24], u102, fmaf(L[23], u111, fmaf(L[22], u120, fmaf(L[21], u201, fmaf(L[20], u210, fmaf(L[19], u300, fmaf(L[48], u004, fmaf(L[47], u013, fmaf(L[46], u022, fmaf(L[45], u031, fmaf(L[44], u040, fmaf(L[43], u103, fmaf(L[42], u112, fmaf(L[41], u121, fmaf(L[40], u130, fmaf(L[39], u202, fmaf(L[38], u211, fmaf(L[37], u220, fmaf(L[36], u301, fmaf(L[35], u310, L[34] * u400))))))))))))))))))))))))))))))))) + L[0];
    acc[1] = fmaf(L[7], uz, fmaf(L[6], uy, fmaf(L[4], ux, fmaf(L[17], u002, fmaf(L[16], u011, fmaf(L[15], u020, fmaf(L[13], u101, fmaf(L[12], u110, fmaf(L[10], u200, fmaf(L[32], u003, fmaf(L[31], u012, fmaf(L[30], u021, fmaf(L[29], u030, fmaf(L[27], u102, fmaf(L[26], u111, fmaf(L[25], u120, fmaf(L[23], u201, fmaf(L[22], u210, fmaf(L[20], u300, fmaf(L[53], u004, fmaf(L[52], u013, fmaf(L[51], u022, fmaf(L[50], u031, fmaf(L[49], u040, fmaf(L[47], u103, fmaf(L[46], u112, fmaf(L[45], u121, fmaf(L[44], u130, fmaf(L[42], u202, fmaf(L[41], u211, fmaf(L[40], u220, fmaf(L[38], u301, fmaf(L[37], u310, L[35] * u400))))))))))))))))))))))))))))))))) + L[1];
    acc[2] = fmaf(L[8], uz, fmaf(L[7], uy, fmaf(L[5], ux, fmaf(L[18], u002, fmaf(L[17], u011, fmaf(L[16], u020, fmaf(L[14], u101, fmaf(L[13], u110, fmaf(L[11], u200, fmaf(L[33], u003, fmaf(L[32], u012, fmaf(L[31], u021, fmaf(L[30], u030, fmaf(L[28], u102, fmaf(L[27], u111, fmaf(L[26], u120, fmaf(L[24], u201, fmaf(L[23], u210, fmaf(L[21], u300, fmaf(L[54], u004, fmaf(L[53], u013, fmaf(L[52], u022, fmaf(L[51], u031, fmaf(L[50], u040, fmaf(L[48], u103, fmaf(L[47], u112, fmaf(L[46], u121, fmaf(L[45], u130, fmaf(L[43], u202, fmaf(L[42], u211, fmaf(L[41], u220, fmaf(L[39], u301, fmaf(L[38], u310, L[36] * u400))))))))))))))))))))))))))))))))) + L[2];
}

}  // namespace hq
