"""Marching-cubes case table used by the dual MC of the jspsr path.

This is the standard 256-case Lorensen-Cline triangle table in the widely published
Bourke corner/edge numbering (corners 0:(0,0,0) 1:(1,0,0) 2:(1,1,0) 3:(0,1,0)
4:(0,0,1) 5:(1,0,1) 6:(1,1,1) 7:(0,1,1)),
the same table csrc/marching_cubes/mc_data.cuh:38-198 holds. Triangle-set parity needs
the identical case table, so the data is the same; the storage is our own: one 60-bit
integer per case, triangle edge ids as little-endian nibbles, 0xF terminates. The edge
mask (edgeTable) and vertex count (numVertsTable) of the reference are derived from it.

TEST INFRASTRUCTURE ONLY (oracle); the HIP library carries its own copy of the same data.
"""

MC_CASES_PACKED = [
0xfffffffffffffff, 0xffffffffffff380, 0xffffffffffff910, 0xfffffffff189381,
0xffffffffffffa21, 0xfffffffffa21380, 0xfffffffff920a29, 0xffffff89a8a2382,
0xffffffffffff2b3, 0xfffffffff0b82b0, 0xfffffffffb32091, 0xffffffb89b912b1,
0xfffffffff3ab1a3, 0xffffffab8a801a0, 0xffffff9ab9b3093, 0xfffffffffb8aa89,
0xffffffffffff874, 0xfffffffff437034, 0xfffffffff748910, 0xffffff137174914,
0xfffffffff748a21, 0xffffffa21403743, 0xffffff748209a29, 0xfff4973727929a2,
0xfffffffff2b3748, 0xffffff40242b74b, 0xffffffb32748109, 0xfff1292b9b49b74,
0xffffff487ab31a3, 0xfff4b7401b41ab1, 0xfff30bab9b09874, 0xffffffab99b4b74,
0xffffffffffff459, 0xfffffffff380459, 0xfffffffff051450, 0xffffff513538458,
0xfffffffff459a21, 0xffffff594a21803, 0xffffff204245a25, 0xfff8434535235a2,
0xfffffffffb32459, 0xffffff594b802b0, 0xffffffb32510450, 0xfff584b82852512,
0xffffff45931ab3a, 0xfffab81a8180594, 0xfff30bab5b05045, 0xffffffb8aa85845,
0xfffffffff975879, 0xffffff375359039, 0xffffff751710870, 0xfffffffff753351,
0xffffff21a759879, 0xfff37503505921a, 0xfff25a758528208, 0xffffff7533525a2,
0xffffff2b3987597, 0xfffb72029279759, 0xfff751871810b32, 0xffffff51771b12b,
0xfffb3a31a758859, 0x0aba010b7905075, 0x07570805a30b0ab, 0xfffffffff5b75ab,
0xffffffffffff56a, 0xfffffffff6a5380, 0xfffffffff6a5109, 0xffffff6a5891381,
0xfffffffff162561, 0xffffff803621561, 0xffffff620609569, 0xfff823625285895,
0xfffffffff56ab32, 0xffffff56a02b80b, 0xffffff6a5b32910, 0xfffb892b92916a5,
0xffffff315356b36, 0xfff6b51505b0b80, 0xfff9505606306b3, 0xffffff89bb96956,
0xfffffffff8746a5, 0xffffffa56374034, 0xffffff7486a5091, 0xfff49737179156a,
0xffffff874156216, 0xfff743403625521, 0xfff620560509748, 0x962695923497937,
0xffffff56a4872b3, 0xfffb720242746a5, 0xfff6a5b32874910, 0x6a54b7b492b9129,
0xfff6b51535b3748, 0xb404b7b016b5b15, 0x74836b630560950, 0xfff9b7974b96956,
0xfffffffffa4694a, 0xffffff380a946a4, 0xffffff04606a10a, 0xfffa16468618138,
0xffffff462421941, 0xfff462942921803, 0xfffffffff624420, 0xffffff624428238,
0xffffff32b46a94a, 0xfff6a4a94b82280, 0xfffa164606102b3, 0x1b8b12184a16146,
0xfff36b319639469, 0x14641916b0181b8, 0xffffff4600636b3, 0xfffffffff86b846,
0xffffffa98a876a7, 0xfffa76a907a0370, 0xfff0818717a176a, 0xffffff37117a76a,
0xfff768981861621, 0x937390976192962, 0xffffff206607087, 0xfffffffff276237,
0xfff76898a86ab32, 0x7a9a76790b72702, 0xb32a767a1871081, 0xfff17616a71b12b,
0x63136b619768698, 0xfffffffff76b190, 0xfff06b0b3607087, 0xffffffffffff6b7,
0xffffffffffffb67, 0xfffffffff67b803, 0xfffffffff67b910, 0xffffff67b138918,
0xfffffffff7b621a, 0xffffff7b6803a21, 0xffffff7b69a2092, 0xfff89a38a3a27b6,
0xfffffffff726327, 0xffffff026067807, 0xffffff910732672, 0xfff678891681261,
0xffffff73171a67a, 0xfff801781a7167a, 0xfff7a69a0a70730, 0xffffff9a88a7a67,
0xfffffffff68b486, 0xffffff640603b63, 0xffffff109648b68, 0xfff63b139369649,
0xffffff1a28b6486, 0xfff640b60b03a21, 0xfff9a2920b648b4, 0x36463b34923a39a,
0xffffff264248328, 0xfffffffff264240, 0xfff834642432091, 0xffffff642241491,
0xfff1a6648168318, 0xffffff40660a01a, 0x39a9303a6834364, 0xfffffffff4a649a,
0xfffffffffb67594, 0xffffff67b594380, 0xffffffb67045105, 0xfff51345343867b,
0xffffffb6721a459, 0xfff594380a217b6, 0xfff204a24a45b67, 0x67b25a523453843,
0xffffff945267327, 0xfff786260680459, 0xfff045051673263, 0x851584812786826,
0xfff73167161a459, 0x459078701671a61, 0xa737a6a305a4a04, 0xfffa84a458a7a67,
0xffffff98b9b6596, 0xfff590650360b63, 0xfffb65510b508b0, 0xffffff1355363b6,
0xfff65b8b9b59a21, 0xa21965690b603b0, 0x52025a50865b58b, 0xfff35a3a25363b6,
0xfff283265825985, 0xffffff260069659, 0x826283865081851, 0xfffffffff612651,
0x698965683a61631, 0xfff06505960a01a, 0xfffffffffa65830, 0xffffffffffff65a,
0xfffffffffb57a5b, 0xffffff03857ba5b, 0xffffff091ba57b5, 0xfff1381897ba57a,
0xffffff15717b21b, 0xfffb27571721380, 0xfff7b2209729579, 0x289823295b27257,
0xffffff573532a52, 0xfff52a578258028, 0xfff2a37353a5109, 0x25752a278129289,
0xfffffffff573531, 0xffffff571170780, 0xffffff735539309, 0xfffffffff795789,
0xffffff8ba8a5485, 0xfff03bba50b5405, 0xfff54aba8a48910, 0x41314943b54a4ba,
0xfff8548b2582152, 0xb151b2b543b0b40, 0x58b8545b2950520, 0xfffffffff3b2549,
0xfff483543253a52, 0xffffff0244252a5, 0x910854583a532a3, 0xfff2492914252a5,
0xffffff153358548, 0xfffffffff501540, 0xfff530509358548, 0xffffffffffff549,
0xffffffba9b947b4, 0xfffba97b9794380, 0xfffb470414b1ba1, 0x4bab474a1843413,
0xfff219b294b97b4, 0x3801b2b197b9479, 0xffffff04224b47b, 0xfff42343824b47b,
0xfff947732972a92, 0x70207872a4797a9, 0xa040a1a472a3a73, 0xfffffffff4782a1,
0xffffff317714194, 0xfff178180714194, 0xfffffffff347304, 0xffffffffffff784,
0xfffffffff8ba8a9, 0xffffffa9bb93903, 0xffffffba88a0a10, 0xfffffffffa3ba13,
0xffffff8b99b1b21, 0xfff9b2921b93903, 0xfffffffffb08b20, 0xffffffffffffb23,
0xffffff98aa82832, 0xfffffffff2902a9, 0xfff8a1810a82832, 0xffffffffffff2a1,
0xfffffffff819831, 0xffffffffffff190, 0xffffffffffff830, 0xfffffffffffffff
]

# Edge e interpolates from corner EDGE_CORNERS[e][0] (p1) to EDGE_CORNERS[e][1] (p2), the argument
# order of the sdf_interp calls in csrc/marching_cubes/dual_mc.cu:100-111.
EDGE_CORNERS = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4),
                (0, 4), (1, 5), (2, 6), (3, 7)]


def case_edges(case: int):
    """Triangle edge list (length multiple of 3) of MC case `case`."""
    v = MC_CASES_PACKED[case]
    out = []
    for k in range(15):
        e = (v >> (4 * k)) & 0xF
        if e == 0xF:
            break
        out.append(e)
    return out


def edge_mask(case: int) -> int:
    m = 0
    for e in case_edges(case):
        m |= 1 << e
    return m


def num_verts(case: int) -> int:
    return len(case_edges(case))
