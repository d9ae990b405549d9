NAME          SC105
ROWS
 N  MAXIM
 L  ROW00001
 L  ROW00002
 L  ROW00003
 E  ROW00004
 E  ROW00005
 E  ROW00006
 E  ROW00007
 L  ROW00008
 L  ROW00009
 L  ROW00010
 L  ROW00011
 L  ROW00012
 L  ROW00013
 E  ROW00014
 E  ROW00015
 E  ROW00016
 E  ROW00017
 E  ROW00018
 L  ROW00019
 L  ROW00020
 L  ROW00021
 L  ROW00022
 L  ROW00023
 L  ROW00024
 E  ROW00025
 E  ROW00026
 E  ROW00027
 E  ROW00028
 E  ROW00029
 L  ROW00030
 L  ROW00031
 L  ROW00032
 L  ROW00033
 L  ROW00034
 L  ROW00035
 E  ROW00036
 E  ROW00037
 E  ROW00038
 E  ROW00039
 E  ROW00040
 L  ROW00041
 L  ROW00042
 L  ROW00043
 L  ROW00044
 L  ROW00045
 L  ROW00046
 E  ROW00047
 E  ROW00048
 E  ROW00049
 E  ROW00050
 E  ROW00051
 L  ROW00052
 L  ROW00053
 L  ROW00054
 L  ROW00055
 L  ROW00056
 L  ROW00057
 E  ROW00058
 E  ROW00059
 E  ROW00060
 E  ROW00061
 E  ROW00062
 L  ROW00063
 L  ROW00064
 L  ROW00065
 L  ROW00066
 L  ROW00067
 L  ROW00068
 E  ROW00069
 E  ROW00070
 E  ROW00071
 E  ROW00072
 E  ROW00073
 L  ROW00074
 L  ROW00075
 L  ROW00076
 L  ROW00077
 L  ROW00078
 L  ROW00079
 E  ROW00080
 E  ROW00081
 E  ROW00082
 E  ROW00083
 E  ROW00084
 L  ROW00085
 L  ROW00086
 L  ROW00087
 L  ROW00088
 L  ROW00089
 L  ROW00090
 E  ROW00091
 E  ROW00092
 E  ROW00093
 E  ROW00094
 E  ROW00095
 L  ROW00096
 L  ROW00097
 L  ROW00098
 L  ROW00099
 L  ROW00100
 L  ROW00101
 E  ROW00102
 L  ROW00103
 L  ROW00104
 L  ROW00105
COLUMNS
    COL00001  ROW00001            2.   ROW00002            1.
    COL00001  ROW00005           -1.
    COL00002  ROW00001            1.   ROW00002            2.
    COL00002  ROW00006           -1.
    COL00003  ROW00001           1.5   ROW00002           1.5
    COL00003  ROW00007           -1.
    COL00004  MAXIM              -1.   ROW00004            1.
    COL00004  ROW00014           1.1
    COL00005  ROW00004           -1.   ROW00013            1.
    COL00005  ROW00015            1.
    COL00006  ROW00005            1.   ROW00008           -1.
    COL00006  ROW00016           -1.
    COL00007  ROW00006            1.   ROW00009           -1.
    COL00007  ROW00017           -1.
    COL00008  ROW00007            1.   ROW00010           -1.
    COL00008  ROW00018           -1.
    COL00009  ROW00008            1.   ROW00011           -1.
    COL00010  ROW00009            1.   ROW00012           -1.
    COL00011  ROW00010            1.   ROW00013           -1.
    COL00012  ROW00011            2.   ROW00012            1.
    COL00012  ROW00016           -1.
    COL00013  ROW00011            1.   ROW00012            2.
    COL00013  ROW00017           -1.
    COL00014  ROW00011           1.5   ROW00012           1.5
    COL00014  ROW00018           -1.
    COL00015  ROW00014           -1.   ROW00015            1.
    COL00015  ROW00025           1.1
    COL00016  ROW00015           -1.   ROW00024            1.
    COL00016  ROW00026            1.
    COL00017  ROW00016            1.   ROW00019           -1.
    COL00017  ROW00027           -1.
    COL00018  ROW00017            1.   ROW00020           -1.
    COL00018  ROW00028           -1.
    COL00019  ROW00018            1.   ROW00021           -1.
    COL00019  ROW00029           -1.
    COL00020  ROW00019            1.   ROW00022           -1.
    COL00021  ROW00020            1.   ROW00023           -1.
    COL00022  ROW00021            1.   ROW00024           -1.
    COL00023  ROW00022            2.   ROW00023            1.
    COL00023  ROW00027           -1.
    COL00024  ROW00022            1.   ROW00023            2.
    COL00024  ROW00028           -1.
    COL00025  ROW00022           1.5   ROW00023           1.5
    COL00025  ROW00029           -1.
    COL00026  ROW00025           -1.   ROW00026            1.
    COL00026  ROW00036           1.1
    COL00027  ROW00026           -1.   ROW00035            1.
    COL00027  ROW00037            1.
    COL00028  ROW00027            1.   ROW00030           -1.
    COL00028  ROW00038           -1.
    COL00029  ROW00028            1.   ROW00031           -1.
    COL00029  ROW00039           -1.
    COL00030  ROW00029            1.   ROW00032           -1.
    COL00030  ROW00040           -1.
    COL00031  ROW00030            1.   ROW00033           -1.
    COL00032  ROW00031            1.   ROW00034           -1.
    COL00033  ROW00032            1.   ROW00035           -1.
    COL00034  ROW00033            2.   ROW00034            1.
    COL00034  ROW00038           -1.
    COL00035  ROW00033            1.   ROW00034            2.
    COL00035  ROW00039           -1.
    COL00036  ROW00033           1.5   ROW00034           1.5
    COL00036  ROW00040           -1.
    COL00037  ROW00036           -1.   ROW00037            1.
    COL00037  ROW00047           1.1
    COL00038  ROW00037           -1.   ROW00046            1.
    COL00038  ROW00048            1.
    COL00039  ROW00038            1.   ROW00041           -1.
    COL00039  ROW00049           -1.
    COL00040  ROW00039            1.   ROW00042           -1.
    COL00040  ROW00050           -1.
    COL00041  ROW00040            1.   ROW00043           -1.
    COL00041  ROW00051           -1.
    COL00042  ROW00041            1.   ROW00044           -1.
    COL00043  ROW00042            1.   ROW00045           -1.
    COL00044  ROW00043            1.   ROW00046           -1.
    COL00045  ROW00044            2.   ROW00045            1.
    COL00045  ROW00049           -1.
    COL00046  ROW00044            1.   ROW00045            2.
    COL00046  ROW00050           -1.
    COL00047  ROW00044           1.5   ROW00045           1.5
    COL00047  ROW00051           -1.
    COL00048  ROW00047           -1.   ROW00048            1.
    COL00048  ROW00058           1.1
    COL00049  ROW00048           -1.   ROW00057            1.
    COL00049  ROW00059            1.
    COL00050  ROW00049            1.   ROW00052           -1.
    COL00050  ROW00060           -1.
    COL00051  ROW00050            1.   ROW00053           -1.
    COL00051  ROW00061           -1.
    COL00052  ROW00051            1.   ROW00054           -1.
    COL00052  ROW00062           -1.
    COL00053  ROW00052            1.   ROW00055           -1.
    COL00054  ROW00053            1.   ROW00056           -1.
    COL00055  ROW00054            1.   ROW00057           -1.
    COL00056  ROW00055            2.   ROW00056            1.
    COL00056  ROW00060           -1.
    COL00057  ROW00055            1.   ROW00056            2.
    COL00057  ROW00061           -1.
    COL00058  ROW00055           1.5   ROW00056           1.5
    COL00058  ROW00062           -1.
    COL00059  ROW00058           -1.   ROW00059            1.
    COL00059  ROW00069           1.1
    COL00060  ROW00059           -1.   ROW00068            1.
    COL00060  ROW00070            1.
    COL00061  ROW00060            1.   ROW00063           -1.
    COL00061  ROW00071           -1.
    COL00062  ROW00061            1.   ROW00064           -1.
    COL00062  ROW00072           -1.
    COL00063  ROW00062            1.   ROW00065           -1.
    COL00063  ROW00073           -1.
    COL00064  ROW00063            1.   ROW00066           -1.
    COL00065  ROW00064            1.   ROW00067           -1.
    COL00066  ROW00065            1.   ROW00068           -1.
    COL00067  ROW00066            2.   ROW00067            1.
    COL00067  ROW00071           -1.
    COL00068  ROW00066            1.   ROW00067            2.
    COL00068  ROW00072           -1.
    COL00069  ROW00066           1.5   ROW00067           1.5
    COL00069  ROW00073           -1.
    COL00070  ROW00069           -1.   ROW00070            1.
    COL00070  ROW00080           1.1
    COL00071  ROW00070           -1.   ROW00079            1.
    COL00071  ROW00081            1.
    COL00072  ROW00071            1.   ROW00074           -1.
    COL00072  ROW00082           -1.
    COL00073  ROW00072            1.   ROW00075           -1.
    COL00073  ROW00083           -1.
    COL00074  ROW00073            1.   ROW00076           -1.
    COL00074  ROW00084           -1.
    COL00075  ROW00074            1.   ROW00077           -1.
    COL00076  ROW00075            1.   ROW00078           -1.
    COL00077  ROW00076            1.   ROW00079           -1.
    COL00078  ROW00077            2.   ROW00078            1.
    COL00078  ROW00082           -1.
    COL00079  ROW00077            1.   ROW00078            2.
    COL00079  ROW00083           -1.
    COL00080  ROW00077           1.5   ROW00078           1.5
    COL00080  ROW00084           -1.
    COL00081  ROW00080           -1.   ROW00081            1.
    COL00081  ROW00091           1.1
    COL00082  ROW00081           -1.   ROW00090            1.
    COL00082  ROW00092            1.
    COL00083  ROW00082            1.   ROW00085           -1.
    COL00083  ROW00093           -1.
    COL00084  ROW00083            1.   ROW00086           -1.
    COL00084  ROW00094           -1.
    COL00085  ROW00084            1.   ROW00087           -1.
    COL00085  ROW00095           -1.
    COL00086  ROW00085            1.   ROW00088           -1.
    COL00087  ROW00086            1.   ROW00089           -1.
    COL00088  ROW00087            1.   ROW00090           -1.
    COL00089  ROW00088            2.   ROW00089            1.
    COL00089  ROW00093           -1.
    COL00090  ROW00088            1.   ROW00089            2.
    COL00090  ROW00094           -1.
    COL00091  ROW00088           1.5   ROW00089           1.5
    COL00091  ROW00095           -1.
    COL00092  ROW00091           -1.   ROW00092            1.
    COL00092  ROW00102           1.1
    COL00093  ROW00092           -1.   ROW00101            1.
    COL00094  ROW00093            1.   ROW00096           -1.
    COL00095  ROW00094            1.   ROW00097           -1.
    COL00096  ROW00095            1.   ROW00098           -1.
    COL00097  ROW00096            1.   ROW00099           -1.
    COL00098  ROW00097            1.   ROW00100           -1.
    COL00099  ROW00098            1.   ROW00101           -1.
    COL00100  ROW00099            2.   ROW00100            1.
    COL00100  ROW00103           -.8   ROW00104            .1
    COL00101  ROW00099            1.   ROW00100            2.
    COL00101  ROW00103            .1   ROW00104           -.8
    COL00102  ROW00099           1.5   ROW00100           1.5
    COL00102  ROW00103           .15   ROW00104           .15
    COL00102  ROW00105           -1.
    COL00103  ROW00102           -1.   ROW00105            1.
RHS
    CONST     ROW00001          200.   ROW00002          100.
    CONST     ROW00011          200.   ROW00012          100.
    CONST     ROW00022          200.   ROW00023          100.
    CONST     ROW00033          200.   ROW00034          100.
    CONST     ROW00044          200.   ROW00045          100.
    CONST     ROW00055          200.   ROW00056          100.
    CONST     ROW00066          200.   ROW00067          100.
    CONST     ROW00077          200.   ROW00078          100.
    CONST     ROW00088          200.   ROW00089          100.
    CONST     ROW00099          200.   ROW00100          100.
ENDATA
